"""Parity of the HIP path (through the C-ABI) with the CPU oracle.

Every stage is run on identical seeded inputs on both sides.  Per-cell and
per-particle results must be bit-identical (cells, offsets, momenta,
interpolator coefficients, J after unload, E and B after the Yee updates);
accumulator sums, whose order differs (the device sums each cell's
contributions in its own fixed order, the oracle in particle order), must
agree within the rounding bound of a reordered sum:
|gpu - oracle| <= 64 eps sum|contributions| per slot.
The device keeps particles binned by cell, so particle sets are compared in
a canonical order (cell, then bit patterns of the members).
"""
import numpy as np
import pytest

import oracle_lib
import states
from cabanapic_amd import Simulation

pytestmark = pytest.mark.gpu

DTYPES = [np.float32, np.float64]
GRIDS = [(6, 5, 4, 8), (1, 7, 1, 16), (9, 1, 3, 5), (3, 3, 3, 12)]


def _name(dtype):
    return "f32" if dtype == np.float32 else "f64"


def sim_from_state(st):
    sim = Simulation(st.nx, st.ny, st.nz, st.np, st.dtype)
    sim.upload_fields(st.fields)
    sim.upload_particles(st.p["x"], st.p["y"], st.p["z"], st.p["ux"], st.p["uy"], st.p["uz"],
                         st.p["w"], st.cell)
    return sim


def assert_particles_bitwise(got, st, members=states.MEMBERS):
    g = states.canonical(got)
    o = states.canonical(states.oracle_particles(st))
    for m in members:
        assert_bitwise(g[m], o[m], m)
    assert_bitwise(g["cell"], o["cell"], "cell")


def assert_bitwise(a, b, what):
    """Equality of the bit patterns (so -0.0 and 0.0 differ)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.shape == b.shape and a.dtype == b.dtype, what
    ui = {1: np.uint8, 4: np.uint32, 8: np.uint64}[a.dtype.itemsize]
    diff = a.view(ui) != b.view(ui)
    n = int(np.count_nonzero(diff))
    first = np.flatnonzero(diff)[:5]
    assert n == 0, f"{what}: {n} of {a.size} differ; first at {first}: {a.ravel()[first]} vs {b.ravel()[first]}"


@pytest.mark.parametrize("dtype", DTYPES, ids=_name)
@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
def test_load_interpolator_bitwise(dtype, grid):
    nx, ny, nz, ppc = grid
    st, _ = states.random_state(nx, ny, nz, 1, dtype, seed=11)
    with sim_from_state(st) as sim:
        sim.load_interpolator_array()
        got = sim.download_interpolators()
    st.load_interpolator()
    assert_bitwise(got, st.interp, "interpolators")


@pytest.mark.parametrize("dtype", DTYPES, ids=_name)
@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
def test_push_particles_bitwise_and_accumulators(dtype, grid):
    nx, ny, nz, ppc = grid
    st, k = states.random_state(nx, ny, nz, ppc, dtype, seed=5 + nx)
    st.load_interpolator()
    with sim_from_state(st) as sim:
        sim.load_interpolator_array()
        sim.clear_accumulator_array()
        sim.push(**k)
        got = sim.download_particles()
        acc = sim.download_accumulators()
        movers = sim.last_push_movers()
        assert sim.device_errors() == 0
    absacc = np.zeros(st.acc.size)
    st.clear_accumulator()
    moved = st.push(**k, abs_acc=absacc)
    assert moved > 0 and movers == moved
    assert_particles_bitwise(got, st)
    eps = np.finfo(dtype).eps
    tol = 64 * eps * absacc.reshape(acc.shape) + np.finfo(dtype).tiny
    err = np.abs(acc.astype(np.float64) - st.acc.astype(np.float64))
    assert (err <= tol).all(), float((err / tol).max())


@pytest.mark.parametrize("dtype", DTYPES, ids=_name)
def test_push_with_random_interpolators(dtype):
    """Interpolators uploaded directly (not derived from fields)."""
    st, k = states.random_state(5, 4, 6, 10, dtype, seed=3)
    rng = np.random.default_rng(4)
    st.interp[:] = rng.uniform(-2, 2, st.interp.shape).astype(dtype)
    with sim_from_state(st) as sim:
        sim.upload_interpolators(st.interp)
        assert_bitwise(sim.download_interpolators(), st.interp, "interp round trip")
        sim.clear_accumulator_array()
        sim.push(**k)
        got = sim.download_particles()
        acc = sim.download_accumulators()
    absacc = np.zeros(st.acc.size)
    st.clear_accumulator()
    st.push(**k, abs_acc=absacc)
    assert_particles_bitwise(got, st)
    tol = 64 * np.finfo(dtype).eps * absacc.reshape(acc.shape) + np.finfo(dtype).tiny
    assert (np.abs(acc.astype(np.float64) - st.acc) <= tol).all()


@pytest.mark.parametrize("dtype", DTYPES, ids=_name)
@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
def test_unload_bitwise(dtype, grid):
    nx, ny, nz, _ = grid
    st, _ = states.random_state(nx, ny, nz, 1, dtype, seed=21)
    rng = np.random.default_rng(22)
    st.acc[:] = rng.uniform(-1, 1, st.acc.shape).astype(dtype)
    dx, dy, dz, dt = dtype(0.3), dtype(0.2), dtype(0.25), dtype(0.07)
    with sim_from_state(st) as sim:
        sim.upload_accumulators(st.acc)
        sim.unload_accumulator_array(dx, dy, dz, dt)
        got = sim.download_fields()
    st.unload(dx, dy, dz, dt)
    assert_bitwise(got, st.fields, "fields after unload")


@pytest.mark.parametrize("dtype", DTYPES, ids=_name)
@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
def test_ghost_update_j_bitwise(dtype, grid):
    nx, ny, nz, _ = grid
    st, _ = states.random_state(nx, ny, nz, 1, dtype, seed=31)
    with sim_from_state(st) as sim:
        sim.update_ghosts_j()
        got = sim.download_fields()
    st.fold_ghosts()
    st.copy_ghosts(6)
    assert_bitwise(got, st.fields, "fields after J ghost update")


@pytest.mark.parametrize("dtype", DTYPES, ids=_name)
@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
def test_advance_b_bitwise(dtype, grid):
    nx, ny, nz, _ = grid
    st, _ = states.random_state(nx, ny, nz, 1, dtype, seed=41)
    p = (dtype(0.21), dtype(0.33), dtype(0.17))
    with sim_from_state(st) as sim:
        sim.advance_b(*p)
        got = sim.download_fields()
    st.advance_b(*p)
    assert_bitwise(got, st.fields, "fields after advance_b")


@pytest.mark.parametrize("dtype", DTYPES, ids=_name)
@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
def test_advance_e_bitwise(dtype, grid):
    nx, ny, nz, _ = grid
    st, _ = states.random_state(nx, ny, nz, 1, dtype, seed=51)
    p = (dtype(0.42), dtype(0.66), dtype(0.34))
    dt_eps0 = dtype(0.09)
    with sim_from_state(st) as sim:
        sim.advance_e(*p, dt_eps0)
        got = sim.download_fields()
    st.advance_e(*p, dt_eps0)
    assert_bitwise(got, st.fields, "fields after advance_e")


@pytest.mark.parametrize("dtype", DTYPES, ids=_name)
def test_energies(dtype):
    st, _ = states.random_state(7, 6, 5, 4, dtype, seed=61)
    with sim_from_state(st) as sim:
        e, b = sim.energies()
        ke = sim.kinetic_energy()
    e0, b0, _, _ = st.energies()
    ke0 = st.kinetic_energy()
    assert abs(e - e0) <= 1e-13 * e0
    assert abs(b - b0) <= 1e-13 * b0
    assert abs(ke - ke0) <= 1e-12 * ke0


@pytest.mark.parametrize("dtype", DTYPES, ids=_name)
@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
def test_uncenter_bitwise(dtype, grid):
    """uncenter_particles (src/uncenter_p.h:4-105) on loaded interpolators:
    momenta bit-identical to the oracle's, positions and cells unchanged."""
    nx, ny, nz, ppc = grid
    st, k = states.random_state(nx, ny, nz, ppc, dtype, seed=61, umax=0.9)
    with sim_from_state(st) as sim:
        sim.load_interpolator_array()
        sim.uncenter_particles(k["qdt_2mc"])
        got = sim.download_particles()
    before = states.oracle_particles(st.copy())
    st.load_interpolator()
    st.uncenter(k["qdt_2mc"])
    assert_particles_bitwise(got, st)
    for m in ("x", "y", "z"):
        assert_bitwise(states.canonical(got)[m], states.canonical(before)[m], m)
    assert not np.array_equal(states.canonical(got)["ux"], states.canonical(before)["ux"])


@pytest.mark.parametrize("dtype", DTYPES, ids=_name)
def test_one_step_particles_bitwise(dtype):
    """A full fused step: the push sees identical interpolators, so the
    particles match bit for bit; fields agree to the reordered-sum bound."""
    nx, ny, nz, ppc = 6, 5, 4, 8
    st, _ = states.random_state(nx, ny, nz, ppc, dtype, seed=71)
    k = states.step_consts(dtype)
    with sim_from_state(st) as sim:
        hist = sim.step(dict(k, energy_interval=1), 1)
        got = sim.download_particles()
        f = sim.download_fields()
    h, _ = st.run(states.consts_vector(k), 1)
    assert_particles_bitwise(got, st)
    scale = np.abs(st.fields).max()
    tol = (1e-5 if dtype == np.float32 else 1e-13) * scale
    assert np.abs(f.astype(np.float64) - st.fields).max() <= tol
    assert abs(hist[0, 0] - h[0, 0]) <= 1e-5 * h[0, 0]
    assert abs(hist[0, 1] - h[0, 1]) <= 1e-5 * h[0, 1]


@pytest.mark.parametrize("dtype", DTYPES, ids=_name)
def test_fused_step_leaves_last_interpolators(dtype):
    """The fused step gathers the interpolators inside the push; after a call
    of n steps the interpolator array holds those of the n-th push, which
    load_interpolator_array computes from the fields after n - 1 steps.
    Without particles the step loads them as a stage of its own."""
    st, _ = states.random_state(7, 5, 4, 6, dtype, seed=83)
    k = states.step_consts(dtype)
    with sim_from_state(st) as a, sim_from_state(st) as b:
        a.step(k, 3, history=False)
        b.step(k, 2, history=False)
        b.load_interpolator_array()
        assert_bitwise(a.download_interpolators(), b.download_interpolators(), "interpolators")
        assert a.device_errors() == 0 and b.device_errors() == 0
    st0, _ = states.random_state(7, 5, 4, 1, dtype, seed=84)
    with Simulation(st0.nx, st0.ny, st0.nz, 1, dtype) as sim:
        sim.upload_fields(st0.fields)
        sim.step(k, 1, history=False)
        got = sim.download_interpolators()
    st0.load_interpolator()
    assert_bitwise(got, st0.interp, "interpolators without particles")


@pytest.mark.parametrize("dtype", DTYPES, ids=_name)
def test_fused_step_equals_stage_calls(dtype):
    """cpic_step (interpolators gathered inside the push, ghost copies written
    by the Yee kernels, no accumulator clear) leaves the same particles,
    fields, accumulators and interpolators, bit for bit, as the reference's
    loop body made of the stage calls (example/example.cpp:216-266)."""
    st, _ = states.random_state(6, 5, 4, 8, dtype, seed=87, umax=0.6)
    k = states.step_consts(dtype)
    with sim_from_state(st) as a, sim_from_state(st) as b:
        a.step(k, 3, history=False)
        for _ in range(3):
            b.load_interpolator_array()
            b.clear_accumulator_array()
            b.push(k["qdt_2mc"], k["cdt_dx"], k["cdt_dy"], k["cdt_dz"], k["qsp"])
            b.unload_accumulator_array(k["dx"], k["dy"], k["dz"], k["dt"])
            b.advance_b(0.5 * k["px"], 0.5 * k["py"], 0.5 * k["pz"])
            b.advance_e(k["px"], k["py"], k["pz"], k["dt_eps0"])
            b.advance_b(0.5 * k["px"], 0.5 * k["py"], 0.5 * k["pz"])
        ga, gb = states.canonical(a.download_particles()), states.canonical(b.download_particles())
        for m in list(states.MEMBERS) + ["cell"]:
            assert_bitwise(ga[m], gb[m], m)
        assert_bitwise(a.download_fields(), b.download_fields(), "fields")
        assert_bitwise(a.download_accumulators(), b.download_accumulators(), "accumulators")
        assert_bitwise(a.download_interpolators(), b.download_interpolators(), "interpolators")
        assert a.device_errors() == 0 and b.device_errors() == 0


@pytest.mark.parametrize("dtype", DTYPES, ids=_name)
def test_many_steps_energy_agreement(dtype):
    """30 fused steps: trajectories drift apart only through the reordered J
    sums; total field energy stays within a small relative band."""
    nx, ny, nz, ppc = 8, 6, 5, 6
    st, _ = states.random_state(nx, ny, nz, ppc, dtype, seed=81, umax=0.3)
    st.fields[:] *= dtype(0.01)
    k = states.step_consts(dtype)
    with sim_from_state(st) as sim:
        hist = sim.step(dict(k, energy_interval=1), 30)
        got = sim.download_particles()
        assert sim.device_errors() == 0
    h, _ = st.run(states.consts_vector(k), 30)
    assert got["cell"].size == st.cell.size
    tot_g = hist[:, 0] + hist[:, 1]
    tot_o = h[:, 0] + h[:, 1]
    rel = np.abs(tot_g - tot_o) / tot_o
    assert rel.max() < (1e-3 if dtype == np.float32 else 1e-9), rel.max()


@pytest.mark.parametrize("dtype", DTYPES, ids=_name)
def test_particles_stay_binned_by_cell(dtype):
    """Downloads list particles cell by cell, before and after pushes."""
    st, k = states.random_state(7, 5, 6, 9, dtype, seed=91)
    with sim_from_state(st) as sim:
        first = sim.download_particles()
        assert np.all(np.diff(first["cell"]) >= 0)
        assert_particles_bitwise(first, st)
        sim.sort_particles()
        sim.load_interpolator_array()
        for _ in range(3):
            sim.push(**k)
        got = sim.download_particles()
    assert np.all(np.diff(got["cell"]) >= 0)
    assert got["cell"].size == st.np


@pytest.mark.parametrize("dtype", DTYPES, ids=_name)
def test_push_is_independent_of_input_order(dtype):
    """Shuffling the uploaded particles changes nothing but summation order."""
    st, k = states.random_state(6, 6, 6, 10, dtype, seed=101)
    st.load_interpolator()
    perm = np.random.default_rng(5).permutation(st.np)
    shuffled = st.copy()
    for m in states.MEMBERS:
        shuffled.p[m] = st.p[m][perm].copy()
    shuffled.cell = st.cell[perm].copy()
    with sim_from_state(shuffled) as sim:
        sim.load_interpolator_array()
        sim.clear_accumulator_array()
        sim.push(**k)
        got = sim.download_particles()
        acc = sim.download_accumulators()
    absacc = np.zeros(st.acc.size)
    st.clear_accumulator()
    st.push(**k, abs_acc=absacc)
    assert_particles_bitwise(got, st)
    tol = 64 * np.finfo(dtype).eps * absacc.reshape(acc.shape) + np.finfo(dtype).tiny
    assert (np.abs(acc.astype(np.float64) - st.acc) <= tol).all()


def _acc_within_reordering_bound(acc, st, absacc, dtype):
    tol = 64 * np.finfo(dtype).eps * absacc.reshape(acc.shape) + np.finfo(dtype).tiny
    err = np.abs(acc.astype(np.float64) - st.acc.astype(np.float64))
    assert (err <= tol).all(), float((err / tol).max())


# (nx, ny, nz, ppc, umax): a small grid, and one whose columns are longer than
# 64 rows (several occupancy and departure words per cell)
REPEAT_CASES = [(6, 5, 4, 8, 0.6), (3, 3, 2, 90, 0.9)]


@pytest.mark.parametrize("dtype", DTYPES, ids=_name)
@pytest.mark.parametrize("case", REPEAT_CASES, ids=lambda c: "x".join(map(str, c[:4])))
def test_repeated_pushes_bitwise(dtype, case):
    """Four pushes on the same interpolators: the first packs the columns,
    the later ones keep every particle in its row (the in-place pass).  The
    particles match the oracle's bit for bit after every push, and the
    accumulators, which every push adds to, agree to the reordered-sum
    bound."""
    nx, ny, nz, ppc, umax = case
    st, k = states.random_state(nx, ny, nz, ppc, dtype, seed=121 + ppc, umax=umax)
    st.load_interpolator()
    st.clear_accumulator()
    absacc = np.zeros(st.acc.size)
    with sim_from_state(st) as sim:
        sim.load_interpolator_array()
        sim.clear_accumulator_array()
        for i in range(4):
            sim.push(**k)
            moved = st.push(**k, abs_acc=absacc)
            assert moved > 0 and sim.last_push_movers() == moved
            assert_particles_bitwise(sim.download_particles(), st)
        acc = sim.download_accumulators()
        assert sim.device_errors() == 0
    _acc_within_reordering_bound(acc, st, absacc, dtype)


@pytest.mark.parametrize("dtype", DTYPES, ids=_name)
def test_crossing_record_overflow_path(dtype):
    """Sixty particles near one corner of a cell move diagonally and cross
    two faces each: the second crossings of one (cell, face) outnumber its
    record list, and the excess takes the overflow list and the spill
    fallback, with identical results."""
    nx, ny, nz = 4, 3, 3
    n = 60
    st = oracle_lib.OracleState(nx, ny, nz, n, dtype)
    rng = np.random.default_rng(7)
    st.cell[:] = 2 + (nx + 2) * (1 + (ny + 2) * 2)  # one interior cell
    st.p["x"][:] = rng.uniform(0.85, 0.99, n).astype(dtype)
    st.p["y"][:] = rng.uniform(0.85, 0.99, n).astype(dtype)
    st.p["z"][:] = rng.uniform(-0.5, 0.5, n).astype(dtype)
    st.p["ux"][:] = rng.uniform(0.95, 1.05, n).astype(dtype)
    st.p["uy"][:] = rng.uniform(0.95, 1.05, n).astype(dtype)
    st.p["uz"][:] = rng.uniform(-0.1, 0.1, n).astype(dtype)
    st.p["w"][:] = dtype(1)
    k = dict(qdt_2mc=dtype(-0.05), cdt_dx=dtype(0.45), cdt_dy=dtype(0.4), cdt_dz=dtype(0.35),
             qsp=dtype(-1.0))
    st.load_interpolator()
    absacc = np.zeros(st.acc.size)
    st.clear_accumulator()
    moved = []
    with sim_from_state(st) as sim:
        sim.load_interpolator_array()
        sim.clear_accumulator_array()
        for _ in range(2):  # a packing push, then an in-place one
            sim.push(**k)
            moved.append(st.push(**k, abs_acc=absacc))
            assert_particles_bitwise(sim.download_particles(), st)
        acc = sim.download_accumulators()
        assert sim.device_errors() == 0
    # every particle crossed the x and the y face of the corner in the first
    # push (33 x first, 27 y first for this seed), far more second crossings
    # per (cell, face) than the 8 record slots of those lists
    assert moved[0] == n
    _acc_within_reordering_bound(acc, st, absacc, dtype)


@pytest.mark.parametrize("dtype", DTYPES, ids=_name)
def test_third_crossing_records_past_the_byte_cap(dtype):
    """1200 particles near one corner of a cell move along the diagonal and
    cross all three faces: the third crossings of some (cell, face) outnumber
    the 248 record slots a byte count allows (the list holds min(248, ...)
    slots), and the excess takes the overflow list and the spill fallback,
    with identical results."""
    nx, ny, nz = 4, 4, 4
    n = 1200
    st = oracle_lib.OracleState(nx, ny, nz, n, dtype)
    rng = np.random.default_rng(11)
    X, Y = nx + 2, ny + 2
    st.cell[:] = 2 + X * (2 + Y * 2)  # interior cell (2, 2, 2)
    for m in ("x", "y", "z"):
        st.p[m][:] = rng.uniform(0.9, 0.99, n).astype(dtype)
    for m in ("ux", "uy", "uz"):
        st.p[m][:] = rng.uniform(1.9, 2.1, n).astype(dtype)
    st.p["w"][:] = dtype(1)
    k = dict(qdt_2mc=dtype(-0.01), cdt_dx=dtype(0.45), cdt_dy=dtype(0.45), cdt_dz=dtype(0.45),
             qsp=dtype(-1.0))
    st.load_interpolator()
    absacc = np.zeros(st.acc.size)
    st.clear_accumulator()
    with sim_from_state(st) as sim:
        sim.load_interpolator_array()
        sim.clear_accumulator_array()
        sim.push(**k)
        st.push(**k, abs_acc=absacc)
        got = sim.download_particles()
        assert_particles_bitwise(got, st)
        acc = sim.download_accumulators()
        assert sim.device_errors() == 0
    # particles that crossed all three faces end in cell (3, 3, 3): at least
    # a third of them share one (cell, face) of their third crossing
    diag = 3 + X * (3 + Y * 3)
    assert np.count_nonzero(st.cell == diag) >= 3 * 249
    _acc_within_reordering_bound(acc, st, absacc, dtype)


def test_invalid_cell_rejected_on_upload():
    st, _ = states.random_state(4, 4, 4, 2, np.float32, seed=1)
    sim = Simulation(4, 4, 4, st.np, np.float32)
    bad = st.cell.copy()
    bad[3] = 0  # a ghost voxel
    with pytest.raises(Exception):
        sim.upload_particles(st.p["x"], st.p["y"], st.p["z"], st.p["ux"], st.p["uy"],
                             st.p["uz"], st.p["w"], bad)
    sim.close()


def test_reflect_boundary_unsupported():
    from cabanapic_amd import CpicError
    st, k = states.random_state(4, 4, 4, 2, np.float32, seed=1)
    with sim_from_state(st) as sim:
        with pytest.raises(CpicError):
            sim.push(**k, boundary=0)


def test_stage_timing_masks():
    """Stage timing counts only the stages asked for: the push alone, then
    every stage of the fused step once per step (the interpolator load runs
    inside the push while particles are resident)."""
    st, _ = states.random_state(4, 3, 3, 6, np.float32, seed=201)
    k = states.step_consts(np.float32)
    with sim_from_state(st) as sim:
        sim.set_timing(True, stages=("push",))
        sim.step(k, 3, history=False)
        t = sim.stage_times()
        assert t["push"][0] == 3 and t["push"][1] > 0
        assert all(v[0] == 0 for name, v in t.items() if name != "push")
        sim.set_timing(True)
        sim.step(dict(k, energy_interval=1), 2, history=True)
        t = sim.stage_times()
        for name in ("push", "unload", "advance_e", "energy"):
            assert t[name][0] == 2, name
        assert t["advance_b"][0] == 4
        assert t["interp"][0] == 0
        sim.set_timing(False)
        sim.step(k, 1, history=False)
        assert all(v[0] == 0 for v in sim.stage_times().values())


# (nx, ny, nz, ppc): the electrostatic solver is 1D (example/example.cpp:69-74)
ES_GRIDS = [(16, 1, 1, 8), (5, 1, 1, 20)]


@pytest.mark.parametrize("dtype", DTYPES, ids=_name)
@pytest.mark.parametrize("grid", ES_GRIDS, ids=lambda g: "x".join(map(str, g)))
def test_es_solver_stages_bitwise(dtype, grid):
    """ES_Field_Solver_1D (src/fields.h:467-545) on random fields with
    nonzero ghost current: advance_b leaves the fields as they are;
    advance_e (the fold with its ghost cells, then E += -dt_eps0 J on every
    voxel) is bit-identical to the oracle; the all-voxel E energy agrees to
    FP64 rounding."""
    nx, ny, nz, ppc = grid
    st, _ = states.random_state(nx, ny, nz, ppc, dtype, seed=211)
    k = states.step_consts(dtype)
    with sim_from_state(st) as sim:
        sim.set_field_solver("es_1d")
        sim.advance_b(k["px"], k["py"], k["pz"])
        assert_bitwise(sim.download_fields(), st.fields, "fields after ES advance_b")
        sim.advance_e(k["px"], k["py"], k["pz"], k["dt_eps0"])
        f = sim.download_fields()
        e, b = sim.energies()
    st.es_advance_e(k["dt_eps0"])
    assert_bitwise(f, st.fields, "fields after ES advance_e")
    e64, _ = st.es_energy()
    b64 = st.energies()[1]
    assert abs(e - e64) <= 1e-12 * e64 and abs(b - b64) <= 1e-12 * b64


@pytest.mark.parametrize("dtype", DTYPES, ids=_name)
def test_es_fused_steps_follow_the_oracle(dtype):
    """cpic_step with the ES solver against the oracle's ES stage loop
    (load_interpolator, push, unload, ES advance_e): particles bit for bit
    after the first step, energies within the reordered-J band over 20
    steps."""
    nx, ny, nz, ppc = 24, 1, 1, 16
    st, _ = states.random_state(nx, ny, nz, ppc, dtype, seed=221, umax=0.3)
    st.fields[:] *= dtype(0.01)
    k = states.step_consts(dtype)
    o = st.copy()
    with sim_from_state(st) as sim:
        sim.set_field_solver("es_1d")
        h1 = sim.step(dict(k, energy_interval=1), 1)
        got = sim.download_particles()
        h = sim.step(dict(k, energy_interval=1), 19)
    hist = np.vstack([h1, h])
    ref = []
    for step in range(20):
        o.load_interpolator()
        o.clear_accumulator()
        o.push(k["qdt_2mc"], k["cdt_dx"], k["cdt_dy"], k["cdt_dz"], k["qsp"])
        o.unload(k["dx"], k["dy"], k["dz"], k["dt"])
        o.es_advance_e(k["dt_eps0"])
        if step == 0:
            assert_particles_bitwise(got, o)
        ref.append(o.es_energy()[0])
    ref = np.array(ref)
    rel = np.abs(hist[:, 0] - ref) / np.abs(ref)
    assert rel.max() < (1e-3 if dtype == np.float32 else 1e-9), rel.max()


def test_es_solver_needs_1d():
    st, _ = states.random_state(4, 3, 1, 2, np.float64, seed=231)
    from cabanapic_amd import CpicError
    with sim_from_state(st) as sim:
        with pytest.raises(CpicError):
            sim.set_field_solver("es_1d")
        sim.set_field_solver("em")
