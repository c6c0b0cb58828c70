"""Pins of the CPU oracle by what the mathematics fixes, independent of the
oracle's own formulas.

The gold energy histories pin the whole step only for the 1D-in-y two-stream
deck (x and z stencils, transverse motion and multi-face crossings are not
exercised there).  These checks close that gap:

* Boris push (src/push.h:124-182): uniform E gives du = 2 qdt_2mc E exactly
  and leaves the other components untouched; uniform B conserves |u|, keeps
  u parallel to B and rotates u about B by -2 qdt_2mc |B| / gamma up to the
  x^7 truncation of the tan series; the rotation sense is that of q u x B.
* uncenter_particles (src/uncenter_p.h:4-105): uniform E with B = 0 adds
  exactly the forward half kick qdt_2mc E (the reference's sign, which VPIC
  reverses); uniform B conserves |u| and u.B and rotates u back by half the
  push's angle, +qdt_2mc |B| / gamma within the tan-series truncation.
* move_p (src/move_p.h:59-227): free streaming advances the unwrapped
  coordinate 2(i-1) + x + 1 by 2 u cdt / gamma per step on every axis,
  through any number of face, edge and corner crossings and periodic wraps;
  small hand-worked crossings land in the expected cell at the expected
  offset.
* deposit + unload + fold (src/push.h:218-256, src/move_p.h:101-209,
  src/accumulator.cpp:45-122, src/fields.h:102-271): discrete continuity of
  trilinear node charge, dQ_node = -dt * (flux of J through the node's dual
  cell), to rounding; periodic translation invariance.
* interpolator (src/interpolator.cpp:4-124): the bilinear (E) and linear
  (B) forms reproduce the edge/face samples at the cell corners.
* Yee solver (src/fields.h:618-719): a standing vacuum mode along each axis
  and each polarization evolves as cos(n theta) with the leapfrog
  dispersion sin(theta/2) = p sin(k d/2); the Gauss residual
  div E - rho / eps0 is invariant under full steps with particles.
* energies (src/fields.h:556-615): brute-force sum in long double.
* electrostatic 1D solver (src/fields.h:467-545): with no ghost current,
  E moves by exactly -dt_eps0 J on every voxel; the Gauss residual is
  invariant over full ES steps with particles; the all-voxel energy is a
  brute-force sum.
"""
import numpy as np
import pytest

import oracle_lib
import states

F64 = np.float64


# ----------------------------------------------------------------------------
# helpers

def _single(nx=4, ny=4, nz=4, np_=1, dtype=F64):
    return oracle_lib.OracleState(nx, ny, nz, np_, dtype)


def _voxel(st, ix, iy, iz):
    return ix + (st.nx + 2) * (iy + (st.ny + 2) * iz)


def _unvoxel(st, c):
    c = np.asarray(c, np.int64)
    ix = c % (st.nx + 2)
    iy = (c // (st.nx + 2)) % (st.ny + 2)
    iz = c // ((st.nx + 2) * (st.ny + 2))
    return ix, iy, iz


def _grid(st, a):
    """(Ng,) -> (nz+2, ny+2, nx+2) view."""
    return a.reshape(st.nz + 2, st.ny + 2, st.nx + 2)


def _interior(st, a):
    return _grid(st, a)[1:-1, 1:-1, 1:-1]


def _push_once(st, k, boundary=1):
    st.load_interpolator()
    st.clear_accumulator()
    return st.push(k["qdt_2mc"], k["cdt_dx"], k["cdt_dy"], k["cdt_dz"], k["qsp"], boundary)


def _uniform_fields(st, e=(0, 0, 0), b=(0, 0, 0)):
    st.fields[:] = 0
    for m in range(3):
        st.fields[m] = e[m]
        st.fields[3 + m] = b[m]


def _node_charge(st, qsp):
    """Trilinear (CIC) node charge: a particle at offset x in [-1, 1] of
    voxel i gives (1 - x)/2 to node i and (1 + x)/2 to node i + 1 (per axis),
    nodes wrapped periodically onto 1..n (returned 0-based)."""
    nx, ny, nz = st.nx, st.ny, st.nz
    ix, iy, iz = _unvoxel(st, st.cell)
    x, y, z = (st.p[m].astype(F64) for m in ("x", "y", "z"))
    q = st.p["w"].astype(F64) * F64(qsp)
    Q = np.zeros((nz, ny, nx))
    for a in (0, 1):
        wx = (1 + (2 * a - 1) * x) / 2
        for b in (0, 1):
            wy = (1 + (2 * b - 1) * y) / 2
            for c in (0, 1):
                wz = (1 + (2 * c - 1) * z) / 2
                np.add.at(Q, ((iz - 1 + c) % nz, (iy - 1 + b) % ny, (ix - 1 + a) % nx),
                          q * wx * wy * wz)
    return Q


def _node_divergence(st, fx, fy, fz, dx, dy, dz):
    """Backward-difference divergence at nodes of edge-centred components
    (x-edge of voxel i joins node i to node i + 1), periodic."""
    ax, ay, az = (_interior(st, f).astype(F64) for f in (fx, fy, fz))
    return ((ax - np.roll(ax, 1, axis=2)) / dx + (ay - np.roll(ay, 1, axis=1)) / dy
            + (az - np.roll(az, 1, axis=0)) / dz)


# ----------------------------------------------------------------------------
# Boris push

@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_uniform_e_kick(dtype):
    rng = np.random.default_rng(1)
    n = 64
    st = _single(np_=n, dtype=dtype)
    E = rng.uniform(-1, 1, 3)
    _uniform_fields(st, e=E)
    st.cell[:] = _voxel(st, 2, 2, 2)
    for m in ("x", "y", "z"):
        st.p[m][:] = 0
    for m in ("ux", "uy", "uz"):
        st.p[m][:] = rng.uniform(-0.5, 0.5, n)
    st.p["w"][:] = 1
    u0 = {m: st.p[m].astype(F64).copy() for m in ("ux", "uy", "uz")}
    qdt_2mc = dtype(-0.03)
    k = dict(qdt_2mc=qdt_2mc, cdt_dx=dtype(1e-3), cdt_dy=dtype(1e-3), cdt_dz=dtype(1e-3),
             qsp=dtype(-1))
    _push_once(st, k)
    eps = np.finfo(dtype).eps
    for c, m in enumerate(("ux", "uy", "uz")):
        want = u0[m] + 2 * F64(qdt_2mc) * F64(dtype(E[c]))
        err = np.abs(st.p[m].astype(F64) - want)
        assert err.max() <= 4 * eps * (np.abs(u0[m]) + 1).max(), (m, err.max())


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_kick_along_one_axis_leaves_others_exact(dtype):
    st = _single(np_=8, dtype=dtype)
    _uniform_fields(st, e=(0.7, 0, 0))
    st.cell[:] = _voxel(st, 1, 3, 2)
    rng = np.random.default_rng(2)
    for m in ("ux", "uy", "uz"):
        st.p[m][:] = rng.uniform(-0.5, 0.5, 8)
    st.p["w"][:] = 1
    uy, uz = st.p["uy"].copy(), st.p["uz"].copy()
    k = dict(qdt_2mc=dtype(0.02), cdt_dx=dtype(1e-3), cdt_dy=dtype(1e-3), cdt_dz=dtype(1e-3),
             qsp=dtype(1))
    _push_once(st, k)
    assert np.array_equal(st.p["uy"], uy) and np.array_equal(st.p["uz"], uz)


@pytest.mark.parametrize("x", [0.02, 0.2])
def test_uniform_b_rotation(x):
    """Rotation of u about B: |u| and u.B conserved, angle -2 qdt_2mc |B|/gamma
    within the tan-series truncation 2 (17/315) x^7 (x = |qdt_2mc B| / gamma).
    A wrong coefficient of the series (1/3 or 2/15) moves the angle by at
    least 2 x^5 / 30, far outside the bound at x = 0.2; a wrong rotation
    sense flips the angle's sign."""
    rng = np.random.default_rng(3)
    n = 32
    bdir = rng.normal(size=3)
    bdir /= np.linalg.norm(bdir)
    u = rng.normal(size=(n, 3)) * 0.4
    gamma = np.sqrt(1 + (u * u).sum(1))
    qdt_2mc = -0.05
    # one |B| per particle so that x is the same for all of them
    bmag = x * gamma / abs(qdt_2mc)
    # a single uniform B is needed per push, so push the particles one by one
    got = np.zeros_like(u)
    for p in range(n):
        one = _single()
        _uniform_fields(one, b=bdir * bmag[p])
        one.cell[:] = _voxel(one, 2, 3, 1)
        one.p["ux"][:], one.p["uy"][:], one.p["uz"][:] = u[p]
        one.p["w"][:] = 1
        k = dict(qdt_2mc=qdt_2mc, cdt_dx=1e-4, cdt_dy=1e-4, cdt_dz=1e-4, qsp=-1.0)
        _push_once(one, k)
        got[p] = one.p["ux"][0], one.p["uy"][0], one.p["uz"][0]
    # interpolated B is exactly the uniform value
    assert np.allclose(np.linalg.norm(got, axis=1), np.linalg.norm(u, axis=1), rtol=1e-14, atol=0)
    assert np.allclose(got @ bdir, u @ bdir, rtol=0, atol=1e-14)
    up = u - np.outer(u @ bdir, bdir)
    gp = got - np.outer(got @ bdir, bdir)
    ang = np.arctan2(np.cross(up, gp) @ bdir, (up * gp).sum(1))
    want = -2 * qdt_2mc * bmag / gamma
    # gamma comes from a float sqrt in the double build: 1.2e-7 relative
    bound = 2 * (17 / 315) * x ** 7 + 1.2e-7 * np.abs(want)
    assert np.all(np.abs(ang - want) <= bound), np.abs(ang - want).max()


# ----------------------------------------------------------------------------
# uncenter_particles

@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_uncenter_uniform_e_adds_forward_half_kick(dtype):
    """With B = 0 the rotation is the identity, so u' = u + qdt_2mc E in
    real_t arithmetic, bit for bit: the sign is the reference's +hax
    (uncenter_p.h:91-93), not the backward kick VPIC applies."""
    rng = np.random.default_rng(11)
    n = 48
    st = _single(np_=n, dtype=dtype)
    E = rng.uniform(-1, 1, 3)
    _uniform_fields(st, e=E)
    st.cell[:] = _voxel(st, 3, 1, 2)
    for m in ("x", "y", "z"):
        st.p[m][:] = rng.uniform(-1, 1, n)
    for m in ("ux", "uy", "uz"):
        st.p[m][:] = rng.uniform(-0.5, 0.5, n)
    u0 = {m: st.p[m].copy() for m in ("ux", "uy", "uz")}
    qdt_2mc = dtype(-0.03)
    st.load_interpolator()
    st.uncenter(qdt_2mc)
    for c, m in enumerate(("ux", "uy", "uz")):
        want = u0[m] + qdt_2mc * dtype(E[c])
        assert want.dtype == dtype and np.array_equal(st.p[m], want), m
    for m in ("x", "y", "z"):
        assert not np.array_equal(st.p[m], np.zeros(n, dtype))


@pytest.mark.parametrize("x", [0.02, 0.2])
def test_uncenter_uniform_b_rotates_back_half_angle(x):
    """With E = 0: |u| and u.B conserved; the angle about B is
    +qdt_2mc |B| / gamma (minus half the push's angle, qdt_4mc = -0.5 qdt_2mc)
    within the tan-series truncation 2 (17/315) x^7 with x = |qdt_4mc B| /
    gamma.  A sign error in qdt_4mc flips the angle; a full qdt_2mc doubles
    it."""
    rng = np.random.default_rng(13)
    n = 24
    bdir = rng.normal(size=3)
    bdir /= np.linalg.norm(bdir)
    u = rng.normal(size=(n, 3)) * 0.4
    gamma = np.sqrt(1 + (u * u).sum(1))
    qdt_2mc = -0.05
    bmag = x * gamma / abs(0.5 * qdt_2mc)
    got = np.zeros_like(u)
    for p in range(n):
        one = _single()
        _uniform_fields(one, b=bdir * bmag[p])
        one.cell[:] = _voxel(one, 1, 2, 3)
        one.p["ux"][:], one.p["uy"][:], one.p["uz"][:] = u[p]
        one.load_interpolator()
        one.uncenter(qdt_2mc)
        got[p] = one.p["ux"][0], one.p["uy"][0], one.p["uz"][0]
    assert np.allclose(np.linalg.norm(got, axis=1), np.linalg.norm(u, axis=1), rtol=1e-14, atol=0)
    assert np.allclose(got @ bdir, u @ bdir, rtol=0, atol=1e-14)
    up = u - np.outer(u @ bdir, bdir)
    gp = got - np.outer(got @ bdir, bdir)
    ang = np.arctan2(np.cross(up, gp) @ bdir, (up * gp).sum(1))
    want = qdt_2mc * bmag / gamma
    bound = 2 * (17 / 315) * x ** 7 + 1e-13
    assert np.all(np.abs(ang - want) <= bound), np.abs(ang - want).max()


# ----------------------------------------------------------------------------
# move_p

def _unwrapped(st):
    """Unwrapped coordinate 2 (i - 1) + (x + 1) per axis, in half cells."""
    ix, iy, iz = _unvoxel(st, st.cell)
    return np.stack([2 * (ix - 1) + st.p["x"].astype(F64) + 1,
                     2 * (iy - 1) + st.p["y"].astype(F64) + 1,
                     2 * (iz - 1) + st.p["z"].astype(F64) + 1], 1)


def test_free_streaming_with_wraps():
    rng = np.random.default_rng(4)
    nx, ny, nz = 5, 4, 3
    n = 400
    st = _single(nx, ny, nz, n)
    st.fields[:] = 0
    c = rng.integers(0, nx * ny * nz, n)
    st.cell[:] = _voxel(st, 1 + c % nx, 1 + (c // nx) % ny, 1 + c // (nx * ny))
    for m in ("x", "y", "z"):
        st.p[m][:] = rng.uniform(-1, 1, n)
    u = rng.uniform(-2, 2, (n, 3))
    # exact face, edge and corner hits from the cell centre
    u[:8] = [[1, 0, 0], [0, -1, 0], [1, 1, 0], [-1, 1, 1], [0, 0, 1], [1, -1, -1], [0, 1, 1], [2, 2, 2]]
    st.p["x"][:8] = st.p["y"][:8] = st.p["z"][:8] = 0
    st.p["ux"][:], st.p["uy"][:], st.p["uz"][:] = u.T
    st.p["w"][:] = 1
    cdt = np.array([0.45, 0.4, 0.35])
    k = dict(qdt_2mc=-0.05, cdt_dx=cdt[0], cdt_dy=cdt[1], cdt_dz=cdt[2], qsp=-1.0)
    g0 = _unwrapped(st)
    inv_gamma = 1 / np.sqrt(1 + (u * u).sum(1))
    period = 2 * np.array([nx, ny, nz])
    steps = 60
    for s in range(1, steps + 1):
        _push_once(st, k)
        # velocity untouched without fields
        assert np.array_equal(np.stack([st.p["ux"], st.p["uy"], st.p["uz"]], 1), u)
        if s % 10:
            continue
        want = g0 + s * 2 * u * cdt * inv_gamma[:, None]
        d = (_unwrapped(st) - want) % period
        d = np.minimum(d, period - d)
        # the double build takes gamma from a float sqrt: 1e-7 relative per step
        tol = s * 2 * np.abs(u * cdt).max() * 2e-7 + 1e-9
        assert d.max() <= tol, (s, d.max())
    ix, iy, iz = _unvoxel(st, st.cell)
    assert ix.min() >= 1 and ix.max() <= nx and iy.min() >= 1 and iy.max() <= ny
    assert iz.min() >= 1 and iz.max() <= nz


def _stream_one(st_shape, cell, pos, u, cdt):
    st = _single(*st_shape)
    st.fields[:] = 0
    st.cell[:] = _voxel(st, *cell)
    st.p["x"][:], st.p["y"][:], st.p["z"][:] = pos
    st.p["ux"][:], st.p["uy"][:], st.p["uz"][:] = u
    st.p["w"][:] = 1
    k = dict(qdt_2mc=0.0, cdt_dx=cdt, cdt_dy=cdt, cdt_dz=cdt, qsp=-1.0)
    moved = _push_once(st, k)
    return st, moved


# u = 0.75 gives gamma = 1.25 (the float sqrt of 1.5625 is exact), so a step
# moves by 2 * 0.75 * cdt / 1.25 = 1.2 cdt half-cells along that axis.
# u = 2^-13 gives gamma = 1 exactly (1 + 2^-26 rounds to 1 in the float sqrt),
# so with cdt = 2048 a step moves by exactly 0.5.
TINY = 2.0 ** -13
@pytest.mark.parametrize("case", [
    dict(cell=(2, 2, 2), pos=(0.5, 0, 0), u=(0.75, 0, 0), cdt=0.5,
         want_cell=(3, 2, 2), want_pos=(-0.9, 0, 0)),           # +x face
    dict(cell=(1, 2, 2), pos=(-0.5, 0, 0), u=(-0.75, 0, 0), cdt=0.5,
         want_cell=(4, 2, 2), want_pos=(0.9, 0, 0)),            # -x face, wrap
    dict(cell=(2, 4, 2), pos=(0, 0.5, 0.5), u=(0, 0.75, 0.75), cdt=0.5,
         want_cell=(2, 1, 3), want_pos=None),                   # y/z tie, y wrap
    dict(cell=(4, 4, 4), pos=(0.5, 0.5, 0.5), u=(0.6, 0.6, 0.6), cdt=1.0,
         want_cell=(1, 1, 1), want_pos=None),                   # corner + 3 wraps
    dict(cell=(2, 2, 2), pos=(0.9, 0.9, 0), u=(0.6, 0.6, 0), cdt=0.2,
         want_cell=(3, 3, 2), want_pos=None),                   # x/y tie: diagonal
    dict(cell=(2, 2, 2), pos=(0.5, 0, 0), u=(TINY, 0, 0), cdt=2048.0,
         want_cell=(2, 2, 2), want_pos=(1.0, 0, 0)),            # lands on the face: stays
    dict(cell=(2, 2, 2), pos=(0, 0, -0.5), u=(0, 0, -TINY), cdt=2048.0,
         want_cell=(2, 2, 2), want_pos=(0, 0, -1.0)),           # same on the -z face
    dict(cell=(2, 2, 2), pos=(0.5 + 2.0 ** -20, 0, 0), u=(TINY, 0, 0), cdt=2048.0,
         want_cell=(3, 2, 2), want_pos=(-1 + 2.0 ** -20, 0, 0)),  # just past the face
])
def test_move_p_hand_cases(case):
    st, moved = _stream_one((4, 4, 4), case["cell"], case["pos"], case["u"], case["cdt"])
    got_cell = tuple(int(v[0]) for v in _unvoxel(st, st.cell))
    assert got_cell == case["want_cell"], got_cell
    assert (moved == 1) == (got_cell != case["cell"])
    u = np.array(case["u"], F64)
    disp = 2 * u * case["cdt"] / np.sqrt(1 + u @ u)
    # offset after crossing: each crossed face shifts the coordinate by -+2
    pos = np.array([st.p["x"][0], st.p["y"][0], st.p["z"][0]], F64)
    shift = 2 * (np.array(got_cell) - np.array(case["cell"]))
    n = np.array([4, 4, 4])
    shift = np.where(shift > 2, shift - 2 * n, np.where(shift < -2, shift + 2 * n, shift))
    want = np.array(case["pos"]) + disp - shift
    assert np.abs(pos - want).max() < 1e-6, (pos, want)
    if case["want_pos"] is not None:
        assert np.abs(pos - np.array(case["want_pos"])).max() < 1e-12
    assert np.all(np.abs(pos) <= 1)


# ----------------------------------------------------------------------------
# deposit, unload and fold

def _deposit_state(seed, nx=5, ny=4, nz=3, ppc=6, umax=1.5):
    st, k = states.random_state(nx, ny, nz, ppc, F64, seed, umax=umax)
    st.fields[:] *= 0.3
    for first in (0, 3):
        st.copy_ghosts(first)
    return st, k


@pytest.mark.parametrize("seed", [5, 6])
def test_discrete_charge_conservation(seed):
    st, k = _deposit_state(seed)
    dt, dx, dy, dz = 0.1, 0.2, 0.25, 0.3
    k["cdt_dx"], k["cdt_dy"], k["cdt_dz"] = dt / dx, dt / dy, dt / dz
    q0 = _node_charge(st, k["qsp"])
    moved = _push_once(st, k)
    assert moved > st.np // 4  # many crossings, some multi-face
    st.fields[6:] = 0
    st.unload(dx, dy, dz, dt)
    st.fold_ghosts()
    q1 = _node_charge(st, k["qsp"])
    f = st.fields
    div_j = _node_divergence(st, f[6], f[7], f[8], dx, dy, dz)
    resid = (q1 - q0) + dt * dx * dy * dz * div_j
    scale = np.abs(st.p["w"]).max() * abs(k["qsp"])
    assert np.abs(q1 - q0).max() > 1e-2 * scale
    assert np.abs(resid).max() < 1e-12 * scale, np.abs(resid).max()


def test_deposit_translation_invariance():
    nx, ny, nz = 5, 4, 3
    st, k = _deposit_state(7, nx, ny, nz)
    sh = st.copy()
    ix, iy, iz = _unvoxel(st, st.cell)
    sh.cell[:] = _voxel(st, 1 + ix % nx, 1 + (iy - 1 + 2) % ny, iz)  # +1 in x, +2 in y
    for m in range(6):
        g = _grid(sh, sh.fields[m])
        g[1:-1, 1:-1, 1:-1] = np.roll(_interior(st, st.fields[m]), (2, 1), axis=(1, 2))
    sh.copy_ghosts(0)
    sh.copy_ghosts(3)
    out = []
    for s in (st, sh):
        _push_once(s, k)
        s.fields[6:] = 0
        s.unload(0.2, 0.25, 0.3, 0.1)
        s.fold_ghosts()
        out.append([_interior(s, s.fields[m]).copy() for m in (6, 7, 8)])
    for a, b in zip(out[0], out[1]):
        assert np.allclose(np.roll(a, (2, 1), axis=(1, 2)), b, rtol=0, atol=1e-13 * np.abs(a).max())


# ----------------------------------------------------------------------------
# interpolator

@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_interpolator_reproduces_corner_samples(dtype):
    nx, ny, nz = 4, 3, 5
    st = _single(nx, ny, nz, 1, dtype)
    st.fields[:] = np.random.default_rng(8).uniform(-1, 1, st.fields.shape).astype(dtype)
    st.load_interpolator()
    I = {n: _interior(st, st.interp[i]).astype(F64) for i, n in enumerate(
        ["ex", "dexdy", "dexdz", "d2exdydz", "ey", "deydz", "deydx", "d2eydzdx",
         "ez", "dezdx", "dezdy", "d2ezdxdy", "cbx", "dcbxdx", "cby", "dcbydy", "cbz", "dcbzdz"])}
    G = [_grid(st, st.fields[m]).astype(F64) for m in range(6)]

    def sample(m, dz_, dy_, dx_):
        return G[m][1 + dz_:nz + 1 + dz_, 1 + dy_:ny + 1 + dy_, 1 + dx_:nx + 1 + dx_]

    tol = 8 * np.finfo(dtype).eps
    for a in (-1, 1):
        for b in (-1, 1):
            ha, hb = (a + 1) // 2, (b + 1) // 2
            # ex over (y, z) = (a, b)
            ex = I["ex"] + a * I["dexdy"] + b * (I["dexdz"] + a * I["d2exdydz"])
            assert np.abs(ex - sample(0, hb, ha, 0)).max() < tol
            # ey over (z, x) = (a, b)
            ey = I["ey"] + a * I["deydz"] + b * (I["deydx"] + a * I["d2eydzdx"])
            assert np.abs(ey - sample(1, ha, 0, hb)).max() < tol
            # ez over (x, y) = (a, b)
            ez = I["ez"] + a * I["dezdx"] + b * (I["dezdy"] + a * I["d2ezdxdy"])
            assert np.abs(ez - sample(2, 0, hb, ha)).max() < tol
        h = (a + 1) // 2
        assert np.abs(I["cbx"] + a * I["dcbxdx"] - sample(3, 0, 0, h)).max() < tol
        assert np.abs(I["cby"] + a * I["dcbydy"] - sample(4, 0, h, 0)).max() < tol
        assert np.abs(I["cbz"] + a * I["dcbzdz"] - sample(5, h, 0, 0)).max() < tol


# ----------------------------------------------------------------------------
# Yee solver

# (propagation axis, E component, B component)
MODES = [(0, 1, 5), (0, 2, 4), (1, 0, 5), (1, 2, 3), (2, 0, 4), (2, 1, 3)]


@pytest.mark.parametrize("axis,ecomp,bcomp", MODES)
def test_yee_standing_wave_dispersion(axis, ecomp, bcomp):
    n_ax, n_tr = 24, 3
    dims = [n_tr, n_tr, n_tr]
    dims[axis] = n_ax
    nx, ny, nz = dims
    st = _single(nx, ny, nz, 1)
    d = [0.3, 0.25, 0.2]
    dt = 0.9 * 0.2 / np.sqrt(3)
    p = [dt / dd for dd in d]
    mode = 3
    kd = 2 * np.pi * mode / n_ax
    idx = np.arange(n_ax + 2) - 1                    # periodic images in the ghosts
    prof = np.sin(kd * idx + 0.37)
    shape = [1, 1, 1]
    shape[2 - axis] = n_ax + 2
    e0 = np.broadcast_to(prof.reshape(shape), (nz + 2, ny + 2, nx + 2)).copy()
    st.fields[:] = 0
    st.fields[ecomp] = e0.ravel()
    st.copy_ghosts(0)
    theta = 2 * np.arcsin(p[axis] * np.sin(kd / 2))
    ref = _interior(st, st.fields[ecomp]).copy()
    nsteps = 40
    for s in range(1, nsteps + 1):
        st.advance_b(0.5 * p[0], 0.5 * p[1], 0.5 * p[2])
        st.advance_e(p[0], p[1], p[2], dt)
        st.advance_b(0.5 * p[0], 0.5 * p[1], 0.5 * p[2])
        got = _interior(st, st.fields[ecomp])
        assert np.abs(got - ref * np.cos(s * theta)).max() < 1e-12 * s, s
    # the other field components stay zero except the paired B
    for m in range(6):
        if m not in (ecomp, bcomp):
            assert np.abs(_interior(st, st.fields[m])).max() < 1e-14
    assert np.abs(_interior(st, st.fields[bcomp])).max() > 0.1


def test_gauss_residual_invariant_over_full_steps():
    nx, ny, nz = 5, 4, 3
    st, _ = _deposit_state(9, nx, ny, nz, ppc=4, umax=1.0)
    st.fields[:] = 0
    k = states.step_consts(F64)
    dx, dy, dz, dt = k["dx"], k["dy"], k["dz"], k["dt"]
    eps0 = dt / k["dt_eps0"]

    def residual():
        f = st.fields
        return (_node_divergence(st, f[0], f[1], f[2], dx, dy, dz)
                - _node_charge(st, k["qsp"]) / (dx * dy * dz * eps0))

    r0 = residual()
    h, moved = st.run(states.consts_vector(k), 25)
    assert moved > 0
    r1 = residual()
    ediv = np.abs(_node_divergence(st, *st.fields[:3], dx, dy, dz)).max()
    assert ediv > 1.0
    assert np.abs(r1 - r0).max() < 1e-11 * ediv, np.abs(r1 - r0).max()


# ----------------------------------------------------------------------------
# energies

@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_energies_brute_force(dtype):
    st = _single(6, 5, 4, 1, dtype)
    st.fields[:] = np.random.default_rng(10).uniform(-2, 2, st.fields.shape).astype(dtype)
    e64, b64, er, br = st.energies()
    ld = np.longdouble
    ew = ld(0.5) * sum((_interior(st, st.fields[m]).astype(ld) ** 2).sum() for m in range(3))
    bw = ld(0.5) * sum((_interior(st, st.fields[m]).astype(ld) ** 2).sum() for m in range(3, 6))
    assert abs(ld(e64) - ew) <= 1e-14 * ew and abs(ld(b64) - bw) <= 1e-14 * bw
    eps = np.finfo(dtype).eps
    assert abs(er - float(ew)) <= 200 * eps * float(ew)
    assert abs(br - float(bw)) <= 200 * eps * float(bw)


# ----------------------------------------------------------------------------
# electrostatic 1D solver

@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_es_advance_e_moves_e_by_minus_dt_j(dtype):
    """With J zero outside the interior the fold changes nothing, and every
    voxel (ghosts included) gets E + (-dt_eps0 J) in real_t; B and J stay."""
    st = _single(7, 1, 1, 1, dtype)
    rng = np.random.default_rng(21)
    st.fields[:] = rng.uniform(-1, 1, st.fields.shape).astype(dtype)
    inner = _grid(st, np.ones(st.ng, bool)).copy()
    inner[:] = False
    inner[1:-1, 1:-1, 1:-1] = True
    for m in (6, 7, 8):
        st.fields[m][~inner.ravel()] = 0
    before = st.fields.copy()
    cj = dtype(0.37)
    st.es_advance_e(cj)
    for m in range(3):
        assert np.array_equal(st.fields[m], before[m] + (-cj) * before[6 + m]), m
    assert np.array_equal(st.fields[3:], before[3:])


def test_es_gauss_residual_invariant_1d():
    """Full electrostatic steps (load_interpolator, push, unload, ES
    advance_e) on a periodic line: div E - rho / eps0 at every node is
    unchanged, which fixes the sign and the dt_eps0 factor of the update
    and the fold it relies on."""
    nx = 9
    rng = np.random.default_rng(22)
    n = 40
    st = _single(nx, 1, 1, n, F64)
    st.cell[:] = _voxel(st, rng.integers(1, nx + 1, n), 1, 1)
    for m in ("x", "y", "z"):
        st.p[m][:] = rng.uniform(-1, 1, n)
    for m in ("ux", "uy", "uz"):
        st.p[m][:] = rng.uniform(-0.8, 0.8, n)
    st.p["w"][:] = rng.uniform(0.5, 1.5, n)
    k = states.step_consts(F64)
    k["dt_eps0"] = k["dt"] / 2.0  # eps0 = 2, so dt_eps0 and dt differ
    dx, dy, dz, dt = k["dx"], k["dy"], k["dz"], k["dt"]
    eps0 = dt / k["dt_eps0"]

    def residual():
        f = st.fields
        return (_node_divergence(st, f[0], f[1], f[2], dx, dy, dz)
                - _node_charge(st, k["qsp"]) / (dx * dy * dz * eps0))

    r0 = residual()
    for _ in range(20):
        _push_once(st, k)
        st.unload(dx, dy, dz, dt)
        st.es_advance_e(k["dt_eps0"])
    r1 = residual()
    ediv = np.abs(_node_divergence(st, *st.fields[:3], dx, dy, dz)).max()
    assert ediv > 1.0
    assert np.abs(r1 - r0).max() < 1e-11 * ediv, np.abs(r1 - r0).max()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_es_energy_brute_force(dtype):
    st = _single(6, 1, 1, 1, dtype)
    st.fields[:] = np.random.default_rng(23).uniform(-2, 2, st.fields.shape).astype(dtype)
    e64, er = st.es_energy()
    ld = np.longdouble
    ew = ld(0.5) * sum((st.fields[m].astype(ld) ** 2).sum() for m in range(3))
    assert abs(ld(e64) - ew) <= 1e-14 * ew
    assert abs(er - float(ew)) <= 200 * np.finfo(dtype).eps * float(ew)
