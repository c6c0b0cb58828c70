"""Pins of the two oracle functions that the gold histories and the physics
checks leave open, and proof that the pins bite.

* ora_kinetic_energy (no reference counterpart; the bench and the decks
  report sum w me c^2 (gamma - 1)): closed forms where gamma is rational,
  |u| = 3/4 -> gamma - 1 = 1/4 and |u| = 4/3 -> gamma - 1 = 2/3, with
  w, me and c all different from 1 so that a dropped factor shows; random
  momenta against gamma - 1 evaluated in long double.
* ora_push_omp, the per-thread duplicate accumulators summed after the push
  (the ScatterView scheme of the reference's OpenMP build,
  example/example.cpp:137-139,248-251): for 1, 2, 7 and 16 threads the
  particles equal the serial ora_push bit for bit and the accumulators agree
  within the reordered-sum bound 64 eps sum|contribution| (the sum over
  threads reorders each slot's terms, nothing else); a second call on fresh
  accumulators gives the same result (the duplicates are reset); ora_run
  with threads agrees with one thread over 50 steps of the 2stream-em deck.
* ora_push_sum's signed sink (the exact-sum reference of the compensated
  device deposit): in the FP64 build it equals the accumulators bit for bit
  (same terms added in the same order in the same type); in the FP32 build,
  with one non-crossing particle per cell (one term per slot) it equals them
  exactly, and on a random state it is within the reordered-sum bound.

The planted-mistake tests build the oracle with an ORA_PLANT_* switch
(a dropped factor, a dropped term, a thread's duplicate left out of the sum,
duplicates not reset) and check that the same pins fail on it.
"""
import numpy as np
import pytest

import oracle_lib
import states

F32, F64 = np.float32, np.float64


# ----------------------------------------------------------------------------
# kinetic energy

def _ke(library, dtype, u, w, me, c):
    u = np.asarray(u, dtype).reshape(-1, 3)
    n = u.shape[0]
    st = oracle_lib.OracleState(1, 1, 1, n, dtype)
    st.library = library
    st.p["ux"][:], st.p["uy"][:], st.p["uz"][:] = u[:, 0], u[:, 1], u[:, 2]
    st.p["w"][:] = np.asarray(w, dtype)
    return st.kinetic_energy(me, c)


def kinetic_failures(library):
    """Names of the failed kinetic-energy pins (empty: all hold)."""
    bad = []
    w, me, c = 2.0, 3.0, 0.5   # exact in float; KE = w me c^2 (gamma - 1)
    # |u| = 3/4 along each axis: gamma = 5/4 exactly
    for dtype in (F32, F64):
        for ax in range(3):
            u = [0.0, 0.0, 0.0]
            u[ax] = 0.75
            got = _ke(library, dtype, u, w, me, c)
            if abs(got - w * me * c * c * 0.25) > 1e-15:
                bad.append(f"u=3/4 axis {ax} {np.dtype(dtype).name}: {got}")
        # |u| = 4/3 split over two axes (0.8, 16/15): gamma = 5/3
        u = [0.8, 0.0, 16.0 / 15.0]
        got = _ke(library, dtype, u, w, me, c)
        uu = np.asarray(u, dtype).astype(np.longdouble)
        ref = w * me * c * c * (np.sqrt(1 + (uu * uu).sum()) - 1)
        if abs(got - float(ref)) > 1e-12 * abs(float(ref)):
            bad.append(f"u=4/3 {np.dtype(dtype).name}: {got} vs {float(ref)}")
        if abs(float(ref) - w * me * c * c * 2 / 3) > 1e-6:
            bad.append("closed form 2/3")
    # sum over many particles against long double, random weights and
    # momenta from non-relativistic to gamma ~ 10
    rng = np.random.default_rng(17)
    u = rng.normal(0, 1, (500, 3)) * rng.choice([0.01, 0.3, 5.0], (500, 1))
    wr = rng.uniform(0.2, 3.0, 500)
    for dtype in (F32, F64):
        ud = u.astype(dtype).astype(np.longdouble)
        wd = wr.astype(dtype).astype(np.longdouble)
        u2 = (ud * ud).sum(axis=1)
        ref = float((wd * u2 / (np.sqrt(1 + u2) + 1)).sum() * np.longdouble(me) * c * c)
        got = _ke(library, dtype, u, wr, me, c)
        if abs(got - ref) > 1e-12 * abs(ref):
            bad.append(f"random {np.dtype(dtype).name}: {got} vs {ref}")
    return bad


def test_kinetic_energy_closed_forms():
    assert kinetic_failures(None) == []


# ----------------------------------------------------------------------------
# OpenMP push

THREADS = (1, 2, 7, 16)


def omp_push_failures(library, dtype, threads=THREADS):
    bad = []
    st0, k = states.random_state(5, 4, 6, 40, dtype, seed=29, umax=1.2)
    st0.library = library
    st0.load_interpolator()
    ref = st0.copy()
    ref.clear_accumulator()
    absacc = np.zeros(ref.acc.size)
    ref.push(**k, abs_acc=absacc)
    tol = 64 * np.finfo(dtype).eps * absacc.reshape(ref.acc.shape) + np.finfo(dtype).tiny
    for nth in threads:
        for call in range(2):  # the second call reuses the persistent duplicates
            st = st0.copy()
            st.clear_accumulator()
            st.push(**k, nthreads=nth)
            for m in states.MEMBERS:
                if not np.array_equal(st.p[m].view(np.uint8), ref.p[m].view(np.uint8)):
                    bad.append(f"{nth} threads call {call}: particle member {m}")
            if not np.array_equal(st.cell, ref.cell):
                bad.append(f"{nth} threads call {call}: cells")
            err = np.abs(st.acc.astype(F64) - ref.acc.astype(F64))
            if not (err <= tol).all():
                bad.append(f"{nth} threads call {call}: accumulators off by "
                           f"{float((err / tol).max()):.3g} x the reordered-sum bound")
    return bad


@pytest.mark.parametrize("dtype", [F32, F64], ids=["f32", "f64"])
def test_push_omp_equals_serial_push(dtype):
    assert omp_push_failures(None, dtype) == []


def omp_run_failures(library, dtype, nsteps=50, threads=(2, 7)):
    """ora_run with threads against one thread on the 2stream-em deck: the
    field energies differ only by the order of the accumulator sums."""
    bad = []
    d, st = oracle_lib.setup_2stream_em(dtype, library)
    one = st.copy()
    h1, m1 = one.run(d["consts"], nsteps, nthreads=1)
    # The deck's early E energy (1e-12 of the beam energy) is made of
    # current noise, so the order of the accumulator sums alone moves it:
    # the serial oracle on three permutations of the particle list differs
    # from the unpermuted run by up to 8 % (E), 1.7 % (B) and 7e-14 (KE) in
    # float and 3.5e-10, 2.2e-11, 4.5e-15 in double over these 50 steps.
    # The bounds are a few times that envelope; a lost or doubled share of
    # the current moves E and B by O(1).
    rtol = {F32: (0.25, 0.05, 1e-12), F64: (1e-8, 1e-9, 1e-13)}[dtype]
    for nth in threads:
        s = st.copy()
        h, m = s.run(d["consts"], nsteps, nthreads=nth)
        if m != m1:
            bad.append(f"{nth} threads: {m} crossings vs {m1}")
        for col, name, tol in zip((0, 1, 4), ("E", "B", "KE"), rtol):
            rel = np.abs(h[:, col] - h1[:, col]) / np.maximum(np.abs(h1[:, col]), 1e-30)
            if rel.max() > tol:
                bad.append(f"{nth} threads: {name} energy off by {rel.max():.3g}")
    return bad


@pytest.mark.parametrize("dtype", [F32, F64], ids=["f32", "f64"])
def test_run_with_threads_equals_one_thread(dtype):
    assert omp_run_failures(None, dtype) == []


# ----------------------------------------------------------------------------
# the pins fail on an oracle with a planted mistake

@pytest.fixture(scope="module")
def plant_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("planted_oracle"))


@pytest.mark.parametrize("plant", ["ORA_PLANT_KE_NO_ME", "ORA_PLANT_KE_GAMMA"])
def test_kinetic_pins_catch_planted_mistake(plant, plant_dir):
    L = oracle_lib.build_planted(plant, plant_dir)
    assert kinetic_failures(L), f"{plant} passed the kinetic-energy pins"


@pytest.mark.parametrize("plant", ["ORA_PLANT_OMP_DROP_DUP", "ORA_PLANT_OMP_NO_RESET"])
def test_omp_pins_catch_planted_mistake(plant, plant_dir):
    L = oracle_lib.build_planted(plant, plant_dir)
    assert omp_push_failures(L, F32, threads=(2, 7)), f"{plant} passed the push pins"
    assert omp_run_failures(L, F64, nsteps=5, threads=(2,)), f"{plant} passed the run pins"


# ----------------------------------------------------------------------------
# the signed-sum sink of the push

def _push_sum(st, k):
    st.load_interpolator()
    st.clear_accumulator()
    a = np.zeros(st.acc.size)
    sm = np.zeros(st.acc.size)
    moved = st.push(**k, abs_acc=a, sum_acc=sm)
    return moved, a.reshape(st.acc.shape), sm.reshape(st.acc.shape)


def test_push_sum_fp64_equals_accumulators():
    st, k = states.random_state(4, 3, 5, 40, F64, seed=81, umax=0.9)
    moved, _, sm = _push_sum(st, k)
    assert moved > 0
    assert np.array_equal(sm, st.acc)


def test_push_sum_fp32_single_term_and_bound():
    st, k = states.random_state(5, 4, 3, 1, F32, seed=82, umax=0.05)
    for m in ("x", "y", "z"):
        st.p[m][:] *= F32(0.5)  # away from the faces: no crossing
    moved, _, sm = _push_sum(st, k)
    assert moved == 0  # one streak per cell: one term per slot
    assert np.array_equal(sm, st.acc.astype(np.float64))
    st, k = states.random_state(4, 4, 4, 100, F32, seed=83, umax=0.9)
    moved, a, sm = _push_sum(st, k)
    assert moved > 0
    err = np.abs(st.acc.astype(np.float64) - sm)
    assert (err <= 64 * np.finfo(F32).eps * a + np.finfo(F32).tiny).all()
    assert err.max() > 0  # the float sums do round
