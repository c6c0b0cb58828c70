"""Pins of the oracle's FP64 reference adjoint (oracle/oracle_adjoint.inc).

* Every transposed field stage against its forward stage by the dot-product
  identity <lambda, A x> = <A^T lambda, x> on random vectors (ghosts
  included): advance_b with its B ghost copy, advance_e with the J fold and
  J ghost copy, unload (fields and accumulators in, fields out) and the
  interpolator load.  A wrong index, sign or coefficient in a transpose
  breaks the identity at O(1).
* The whole step (push, move_p and deposit included) against central
  differences of the smooth FP64 oracle (exact sqrt, so the step is a smooth
  function of its inputs away from crossing decisions): for random
  directions in particles and fields and a random linear objective of the
  state after 1 and 3 steps, <lambda_0, delta> equals the difference
  quotient to 1e-7.
* The faithful FP64 adjoint (the reference's sqrtf narrowing in the primal)
  agrees with the smooth one to the float resolution of gamma.
"""
import ctypes

import numpy as np
import pytest

import oracle_lib
import states

F64 = np.float64


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _lib_fn(name, suffix="_d"):
    return getattr(oracle_lib.lib(), name + suffix)


def _rand(shape, seed):
    return np.random.default_rng(seed).uniform(-1, 1, shape)


GRID = (4, 3, 5)


def _ng(nx, ny, nz):
    return (nx + 2) * (ny + 2) * (nz + 2)


@pytest.mark.parametrize("which", ["advance_b", "advance_e"])
def test_field_stage_transposes_dot_product(which):
    nx, ny, nz = GRID
    ng = _ng(nx, ny, nz)
    px, py, pz, cj = 0.31, 0.27, 0.23, 0.17
    x = _rand((9, ng), 1)
    lam = _rand((9, ng), 2)
    y = x.copy()
    if which == "advance_b":
        _lib_fn("ora_advance_b")(_p(y), px, py, pz, nx, ny, nz)
        code = 0
    else:
        _lib_fn("ora_advance_e")(_p(y), px, py, pz, nx, ny, nz, cj)
        code = 1
    lt = lam.copy()
    _lib_fn("ora_stage_adjoint")(code, _p(lt), None, None, nx, ny, nz, px, py, pz, cj, 0.0, 0.0)
    lhs = float(np.vdot(lam, y))
    rhs = float(np.vdot(lt, x))
    assert abs(lhs - rhs) <= 1e-13 * (abs(lhs) + np.abs(lam).sum()), (lhs, rhs)


def test_unload_transpose_dot_product():
    nx, ny, nz = GRID
    ng = _ng(nx, ny, nz)
    dx, dy, dz, dt = 0.3, 0.2, 0.25, 0.07
    cx, cy, cz = 0.25 / (dy * dz * dt), 0.25 / (dz * dx * dt), 0.25 / (dx * dy * dt)
    f = _rand((9, ng), 3)
    acc = _rand((ng, 3, 4), 4)
    lam = _rand((9, ng), 5)
    y = f.copy()
    _lib_fn("ora_unload_accumulator")(_p(y), _p(acc), nx, ny, nz, dx, dy, dz, dt)
    lt = lam.copy()
    lacc = np.zeros((ng, 3, 4))
    _lib_fn("ora_stage_adjoint")(2, _p(lt), _p(lacc), None, nx, ny, nz, 0.0, 0.0, 0.0, cx, cy, cz)
    lhs = float(np.vdot(lam, y))
    rhs = float(np.vdot(lt, f) + np.vdot(lacc, acc))
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + 1), (lhs, rhs)


def test_load_interpolator_transpose_dot_product():
    nx, ny, nz = GRID
    ng = _ng(nx, ny, nz)
    f = _rand((9, ng), 6)
    li = _rand((18, ng), 7)
    interp = np.zeros((18, ng))
    _lib_fn("ora_load_interpolator")(_p(f), _p(interp), nx, ny, nz)
    lt = np.zeros((9, ng))
    _lib_fn("ora_stage_adjoint")(3, _p(lt), None, _p(li), nx, ny, nz, 0.0, 0.0, 0.0, 0, 0, 0)
    lhs = float(np.vdot(li, interp))
    rhs = float(np.vdot(lt, f))
    assert abs(lhs - rhs) <= 1e-13 * (abs(lhs) + np.abs(li).sum()), (lhs, rhs)


def _state(smooth, seed=31):
    st, _ = states.random_state(*GRID, 6, F64, seed=seed, umax=0.8)
    st.p["w"][:] *= 0.05  # currents small enough for a few linear steps
    st.smooth = smooth
    k = states.step_consts(F64)
    return st, states.consts_vector(k)


def _pack(st):
    """The state vector: fields (9 x Ng) then x, y, z, ux, uy, uz."""
    return np.concatenate([st.fields.ravel()] + [st.p[m] for m in ("x", "y", "z", "ux", "uy",
                                                                    "uz")])


def _unpack_into(st, v):
    ng9 = st.fields.size
    st.fields[:] = v[:ng9].reshape(st.fields.shape)
    n = st.np
    for i, m in enumerate(("x", "y", "z", "ux", "uy", "uz")):
        st.p[m][:] = v[ng9 + i * n: ng9 + (i + 1) * n]


def _run(st, consts, nsteps):
    s = st.copy()
    s.run(consts, nsteps, hist=False)
    return s


def _adjoint(st, consts, nsteps, c):
    """<c, state after nsteps> differentiated: lambda of the initial state."""
    traj = [st.copy()]
    for _ in range(nsteps):
        traj.append(_run(traj[-1], consts, 1))
    ng9 = st.fields.size
    lf = c[:ng9].reshape(9, -1).copy()  # updated in place by step_adjoint
    lp = c[ng9:].reshape(6, -1).copy()
    for s in reversed(traj[:-1]):
        lp = s.step_adjoint(consts, lf, lp)
    return np.concatenate([lf.ravel(), lp.ravel()])


@pytest.mark.parametrize("nsteps", [1, 3])
def test_step_adjoint_against_central_differences(nsteps):
    st, consts = _state(smooth=True)
    v0 = _pack(st)
    rng = np.random.default_rng(8)
    c = rng.uniform(-1, 1, v0.size)
    lam = _adjoint(st, consts, nsteps, c)
    eps = 1e-6
    for trial in range(3):
        d = rng.uniform(-1, 1, v0.size)
        vals = []
        for sgn in (1, -1):
            s = st.copy()
            _unpack_into(s, v0 + sgn * eps * d)
            vals.append(float(np.vdot(c, _pack(_run(s, consts, nsteps)))))
        fd = (vals[0] - vals[1]) / (2 * eps)
        ad = float(np.vdot(lam, d))
        print(f"{nsteps} steps, direction {trial}: adjoint {ad:.12e}, central differences {fd:.12e}")
        assert abs(ad - fd) <= 1e-7 * max(abs(fd), 1.0), (ad, fd)


def test_faithful_adjoint_close_to_smooth():
    """The faithful FP64 step takes gamma through sqrtf (float resolution);
    its adjoint is the smooth one's up to that resolution."""
    st_s, consts = _state(smooth=True)
    st_f, _ = _state(smooth=False)
    rng = np.random.default_rng(9)
    c = rng.uniform(-1, 1, _pack(st_s).size)
    ls = _adjoint(st_s, consts, 2, c)
    lfa = _adjoint(st_f, consts, 2, c)
    rel = np.abs(ls - lfa).max() / np.abs(ls).max()
    print(f"faithful vs smooth adjoint: max rel. difference {rel:.2e}")
    assert rel < 1e-5
