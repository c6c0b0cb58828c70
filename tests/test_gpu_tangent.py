"""Forward-mode derivative of the step (cpic_tangent_* through the C-ABI).

The reference has no derivative of its step, so there is no reference
vector to compare with.  The primal part of the tangent path is held to the
oracle bit for bit (particles) and to the reordered-sum bound
(accumulators).  The tangent part is held to central finite differences of
the oracle's FP64 step on the same inputs: d/dtheta of every particle
coordinate and accumulator after one push, and of the field energies over
several full steps, for theta a particle momentum drift, a guide field B_z,
and random directions in particle and field space.

The differences are taken on the smooth FP64 instantiation of the oracle
(double sqrt for the reference's sqrtf).  The reference-faithful double
build computes gamma through sqrtf, a staircase at float resolution: its
central differences with eps = 1e-6 and 1e-5 differ by about 3e-2 in x after
one push, against 2e-10 for the smooth build.  The GPU tangent is the
derivative of the smooth model evaluated on the faithful trajectory (the two
trajectories differ by about 3e-9 relative).
"""
import numpy as np
import pytest

import states
from cabanapic_amd import TangentSimulation

pytestmark = pytest.mark.gpu

DTYPES = [np.float32, np.float64]
MEMBERS = states.MEMBERS
TMEMBERS = ("x", "y", "z", "ux", "uy", "uz")


def _name(dtype):
    return "f32" if dtype == np.float32 else "f64"


def tangent_from_state(st, tfields=None, tpart=None):
    ts = TangentSimulation(st.nx, st.ny, st.nz, st.np, st.dtype)
    ts.primal.upload_fields(st.fields)
    ts.tangent.upload_fields(np.zeros((9, st.ng)) if tfields is None else tfields)
    t = np.zeros((6, st.np)) if tpart is None else tpart
    ts.upload_particles(*[st.p[m] for m in MEMBERS], st.cell, t)
    return ts


def perturbed(st, eps, tfields, tpart):
    """The state st + eps * (tfields, tpart), in FP64, on the smooth oracle."""
    s = st.copy()
    s.smooth = True
    s.fields += eps * tfields
    for k, m in enumerate(TMEMBERS):
        s.p[m] += eps * tpart[k]
    return s


def assert_same_bits(a, b, what):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.shape == b.shape and a.dtype == b.dtype, what
    ui = {4: np.uint32, 8: np.uint64}[a.dtype.itemsize]
    n = int(np.count_nonzero(a.view(ui) != b.view(ui)))
    assert n == 0, f"{what}: {n} of {a.size} differ"


def sort_by_cell(st):
    order = np.argsort(st.cell, kind="stable")
    st.cell = np.ascontiguousarray(st.cell[order])
    st.p = {m: np.ascontiguousarray(v[order]) for m, v in st.p.items()}


@pytest.mark.parametrize("dtype", DTYPES, ids=_name)
@pytest.mark.parametrize("order", ["shuffled", "sorted"])
def test_tangent_push_primal_bitwise(dtype, order):
    """One push of the tangent path with zero tangents: particles equal the
    oracle's push bit for bit, in upload order, and the movers count with
    them; accumulators agree to the reordered-sum bound; every tangent
    stays exactly zero.  Particles in random cell order, and sorted by cell
    (the deposits of a cell are then summed over runs of up to 24 lanes of a
    wave before the atomics)."""
    st, k = states.random_state(6, 5, 4, 24, dtype, seed=131, umax=0.9)
    if order == "sorted":
        sort_by_cell(st)
    with tangent_from_state(st) as ts:
        for sim in (ts.primal, ts.tangent):
            sim.load_interpolator_array()
            sim.clear_accumulator_array()
        ts.push(**k)
        got = ts.download_particles()
        acc = ts.primal.download_accumulators()
        tacc = ts.tangent.download_accumulators()
        movers = ts.last_push_movers()
        assert ts.primal.device_errors() == 0
    st.load_interpolator()
    st.clear_accumulator()
    absacc = np.zeros(st.acc.size)
    moved = st.push(**k, abs_acc=absacc)
    for m in MEMBERS:
        assert_same_bits(got[m], st.p[m], m)
    assert_same_bits(got["cell"], st.cell, "cell")
    assert moved > 0 and movers == moved
    tol = 64 * np.finfo(dtype).eps * absacc.reshape(acc.shape) + np.finfo(dtype).tiny
    assert (np.abs(acc.astype(np.float64) - st.acc.astype(np.float64)) <= tol).all()
    assert not got["tangent"].any() and not tacc.any()


def test_tangent_push_matches_central_differences():
    """d/dtheta of one push along a random direction of particle coordinates,
    momenta and fields: the tangents of every particle and accumulator equal
    central differences of the smooth FP64 oracle's push (eps = 1e-6), to
    1e-6 relative."""
    st, k = states.random_state(5, 4, 4, 10, np.float64, seed=141, umax=0.9)
    sort_by_cell(st)
    rng = np.random.default_rng(142)
    tfields = rng.uniform(-1, 1, st.fields.shape)
    tpart = rng.uniform(-1, 1, (6, st.np))
    with tangent_from_state(st, tfields, tpart) as ts:
        for sim in (ts.primal, ts.tangent):
            sim.load_interpolator_array()
            sim.clear_accumulator_array()
        ts.push(**k)
        got = ts.download_particles()
        tacc = ts.tangent.download_accumulators()
        assert ts.primal.device_errors() == 0
    eps = 1e-6
    runs = []
    for sgn in (1, -1):
        s = perturbed(st, sgn * eps, tfields, tpart)
        s.load_interpolator()
        s.clear_accumulator()
        s.push(**k)
        runs.append(s)
    plus, minus = runs
    # crossings are taken as they fall on both sides
    assert (plus.cell == minus.cell).all() and (plus.cell == got["cell"]).all()
    for i, m in enumerate(TMEMBERS):
        fd = (plus.p[m] - minus.p[m]) / (2 * eps)
        err = np.abs(got["tangent"][i] - fd)
        assert (err <= 1e-6 * (1 + np.abs(fd))).all(), (m, float(err.max()))
    fd = (plus.acc - minus.acc) / (2 * eps)
    err = np.abs(tacc - fd)
    assert (err <= 1e-6 * (1 + np.abs(fd).max())).all(), float(err.max())
    assert np.abs(tacc).max() > 0.1


def gentle_state(dtype, seed):
    """A random state whose plasma frequency dt resolves: fields and
    weights scaled down (with unit weights the 8 particles per 0.2-wide cell
    make the energies grow twentyfold in ten steps, and central differences
    no longer converge as eps shrinks)."""
    st, _ = states.random_state(6, 5, 4, 8, dtype, seed=seed, umax=0.3)
    st.fields[:] *= dtype(0.01)
    st.p["w"][:] *= dtype(0.01)
    return st


def _direction(st, which):
    """Initial tangents of the deck parameters: a drift of ux (every
    particle's d ux / d theta = 1), or a uniform guide field B_z."""
    tfields = np.zeros(st.fields.shape)
    tpart = np.zeros((6, st.np))
    if which == "drift":
        tpart[3] = 1.0
    else:
        tfields[5] = 1.0
    return tfields, tpart


@pytest.mark.parametrize("which", ["drift", "bz"])
def test_tangent_steps_match_central_differences(which):
    """Ten full steps (interpolator load, push, unload, advance_b/e): the
    derivatives of the E and B energies at every step, and of every particle
    coordinate at the end, equal central differences of the smooth FP64
    oracle's step loop (eps = 1e-6); the primal energies equal the faithful
    oracle's to the reordered-sum level."""
    nsteps = 10
    st = gentle_state(np.float64, 151)
    k = states.step_consts(np.float64)
    tfields, tpart = _direction(st, which)
    with tangent_from_state(st, tfields, tpart) as ts:
        hist = ts.step(k, nsteps)
        got = ts.download_particles()
        assert ts.primal.device_errors() == 0
    h0, _ = st.copy().run(states.consts_vector(k), nsteps)
    assert np.abs(hist[:, 0] - h0[:, 0]).max() <= 1e-9 * np.abs(h0[:, 0]).max()
    assert np.abs(hist[:, 1] - h0[:, 1]).max() <= 1e-9 * np.abs(h0[:, 1]).max()
    eps = 1e-6
    runs = []
    for sgn in (1, -1):
        s = perturbed(st, sgn * eps, tfields, tpart)
        h, _ = s.run(states.consts_vector(k), nsteps)
        runs.append((s, h))
    (plus, hp), (minus, hm) = runs
    for col, what in ((0, "E"), (1, "B")):
        fd = (hp[:, col] - hm[:, col]) / (2 * eps)
        err = np.abs(hist[:, 2 + col] - fd)
        assert err.max() <= 1e-5 * np.abs(fd).max(), (what, float(err.max()), float(np.abs(fd).max()))
    same = (plus.cell == minus.cell) & (plus.cell == got["cell"])
    assert same.mean() > 0.99
    for i, m in enumerate(TMEMBERS):
        fd = (plus.p[m] - minus.p[m]) / (2 * eps)
        err = np.abs(got["tangent"][i] - fd)[same]
        assert (err <= 1e-5 * (1 + np.abs(fd[same]))).all(), (m, float(err.max()))


def test_tangent_float_follows_double():
    """The FP32 step with FP64 tangents, from the same initial state as the
    FP64 step: the energy derivatives agree to 1e-3 relative over ten steps."""
    nsteps = 10
    s32 = gentle_state(np.float32, 161)
    s64 = s32.copy()
    s64.dtype = np.dtype(np.float64)
    s64.fields = s32.fields.astype(np.float64)
    s64.p = {m: v.astype(np.float64) for m, v in s32.p.items()}
    out = []
    for st, dt in ((s32, np.float32), (s64, np.float64)):
        tfields, tpart = _direction(st, "drift")
        with tangent_from_state(st, tfields, tpart) as ts:
            out.append(ts.step(states.step_consts(dt), nsteps))
    h32, h64 = out
    for col in (2, 3):
        scale = np.abs(h64[:, col]).max()
        assert scale > 0
        assert np.abs(h32[:, col] - h64[:, col]).max() <= 1e-3 * scale, col


def test_tangent_rejects_bad_input():
    """Upload checks: interior cells, tangent array present; the reference's
    Reflect boundary is unsupported, as for cpic_push."""
    from cabanapic_amd import CpicError
    st, k = states.random_state(3, 3, 3, 2, np.float64, seed=171)
    with TangentSimulation(3, 3, 3, st.np, np.float64) as ts:
        bad = st.cell.copy()
        bad[0] = 0  # a ghost voxel
        with pytest.raises(CpicError):
            ts.upload_particles(*[st.p[m] for m in MEMBERS], bad, np.zeros((6, st.np)))
        ts.upload_particles(*[st.p[m] for m in MEMBERS], st.cell, np.zeros((6, st.np)))
        with pytest.raises(CpicError):
            ts.push(**k, boundary=0)


def test_tangent_2stream_deck_growth_phase():
    """The reference's 2stream-em regression deck (1D in y, 32 cells x 100
    particles, tests/energy_comparison/2stream-em.cxx): derivatives of the E
    and B energies with respect to the beam drift v0 over the first 300 steps
    of the two-stream instability.  The initial tangent is d(ux)/d(v0) of the
    deck's initializer (only ux depends on v0); the tangent step agrees with
    central differences of the smooth FP64 oracle (eps = 1e-7) to 1e-3 of
    the largest derivative.  The bound is set by the two trajectories, not
    by the derivative: the tangent follows the reference's trajectory (gamma
    through sqrtf), the differences follow the smooth oracle's, and in this
    unstable phase the two runs' energies themselves drift apart by up to
    2.1e-4 (E) and 1.1e-4 (B) of their maximum over the 300 steps (the
    smooth differences at eps = 1e-7 and 1e-6 agree to 3e-7)."""
    import oracle_lib
    nsteps = 300
    d, st = oracle_lib.setup_2stream_em(np.float64)
    v0 = d["v0"]

    def init(v):
        s = oracle_lib.OracleState(d["nx"], d["ny"], d["nz"], d["np"], np.float64)
        oracle_lib.default_particle_init(s, d["dxp"], d["nppc"], d["we"], v)
        return s

    a, b = init(v0 * (1 + 1e-4)), init(v0 * (1 - 1e-4))
    assert (a.cell == b.cell).all()
    assert all(np.array_equal(a.p[m], b.p[m]) for m in ("x", "y", "z", "uy", "uz", "w"))
    tpart = np.zeros((6, st.np))
    tpart[3] = (a.p["ux"] - b.p["ux"]) / (2e-4 * v0)
    names = ["qdt_2mc", "cdt_dx", "cdt_dy", "cdt_dz", "qsp", "dx", "dy", "dz", "dt",
             "px", "py", "pz", "dt_eps0", "me", "c"]
    k = dict(zip(names, (float(v) for v in d["consts"])))
    with tangent_from_state(st, None, tpart) as ts:
        hist = ts.step(k, nsteps)
        assert ts.primal.device_errors() == 0
    eps = 1e-7
    h = []
    for sgn in (1, -1):
        s = perturbed(st, sgn * eps, np.zeros(st.fields.shape), tpart)
        hh, _ = s.run(d["consts"], nsteps)
        h.append(hh)
    for col, what in ((0, "E"), (1, "B")):
        fd = (h[0][:, col] - h[1][:, col]) / (2 * eps)
        err = np.abs(hist[:, 2 + col] - fd)
        assert err.max() <= 1e-3 * np.abs(fd).max(), (what, float(err.max()), float(np.abs(fd).max()))


def test_tangent_graph_replays_match_eager_launches():
    """The tangent step replayed as a HIP graph (the default) against the
    same steps launched stage by stage: chunks of 4, 4 and 1 steps, then 6
    (a longer history, so a new buffer and a new capture), with downloads
    and synchronizations between them, each checked against the eager run.
    The deposits are atomic, so the two runs agree to the reordered-sum level:
    energies and derivatives within 1e-10, particles within 1e-9."""
    st = gentle_state(np.float64, 181)
    k = states.step_consts(np.float64)
    tfields, tpart = _direction(st, "drift")
    runs = {}
    for graph in (True, False):
        with tangent_from_state(st, tfields, tpart) as ts:
            ts.set_graph(graph)
            chunks = []
            for n in (4, 4, 1, 6):
                chunks.append((ts.step(k, n), ts.download_particles()))
                assert ts.primal.device_errors() == 0
            runs[graph] = chunks
    for i, ((hg, pg), (he, pe)) in enumerate(zip(runs[True], runs[False])):
        assert np.isfinite(hg).all(), i
        for col in range(4):
            scale = np.abs(he[:, col]).max()
            assert np.abs(hg[:, col] - he[:, col]).max() <= 1e-10 * scale, (i, col)
        assert np.array_equal(pg["cell"], pe["cell"]), i
        for m in MEMBERS:
            assert np.abs(pg[m] - pe[m]).max() <= 1e-9 * (1 + np.abs(pe[m]).max()), (i, m)
        assert np.abs(pg["tangent"] - pe["tangent"]).max() <= 1e-9 * (
            1 + np.abs(pe["tangent"]).max()), i
