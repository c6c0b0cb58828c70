"""Column capacity under bunching (cpic_config.column_rows, automatic).

Weightless particles (no charge, so the fields stay exactly zero and every
push is free streaming) converge on one plane: the particle at distance d
from the focus moves d / T per step towards it, so the count of the focus
cells grows as 1 / (1 - t/T).  The columns sized at upload (64 rows for 8
per cell) must grow on the way; after the run every particle must equal the
oracle's free-streaming run bit for bit, with no particle lost.
"""
import time

import numpy as np
import pytest

import oracle_lib
import states
from cabanapic_amd import Simulation

pytestmark = pytest.mark.gpu

NX, NY, NZ, PPC = 64, 4, 4, 8
T = 128  # steps to the focus
STEPS = 119  # 0.93 T: the focus cells hold about 14x their initial count


def focusing_state(dtype):
    rng = np.random.default_rng(17)
    nc = NX * NY * NZ
    n = nc * PPC
    st = oracle_lib.OracleState(NX, NY, NZ, n, dtype)
    st.fields[:] = 0
    c = np.repeat(np.arange(nc), PPC)
    ix, iy, iz = 1 + c % NX, 1 + (c // NX) % NY, 1 + c // (NX * NY)
    st.cell[:] = ix + (NX + 2) * (iy + (NY + 2) * iz)
    for m in ("x", "y", "z"):
        st.p[m][:] = rng.uniform(-1, 1, n).astype(dtype)
    # cell widths from the focus plane x = NX/2 (in cells), velocity d / (T cdt)
    xc = (ix - 1) + 0.5 * (st.p["x"].astype(np.float64) + 1.0)
    d = NX / 2 - xc
    cdt = 0.5
    v = d / (T * cdt)
    st.p["ux"][:] = (v / np.sqrt(1.0 - v * v)).astype(dtype)
    st.p["uy"][:] = 0
    st.p["uz"][:] = 0
    st.p["w"][:] = 0
    k = states.step_consts(dtype, cdt=(cdt, cdt, cdt), dt=0.5)
    return st, k


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_columns_grow_under_bunching(dtype):
    st, k = focusing_state(dtype)
    counts0 = np.bincount(st.cell).max()
    with Simulation(NX, NY, NZ, st.np, dtype) as sim:
        sim.upload_particles(*[st.p[m] for m in states.MEMBERS], st.cell)
        rows0 = sim.column_rows()
        sim.step(dict(k, energy_interval=0), STEPS)
        assert sim.device_errors() == 0
        got = sim.download_particles()
        f = sim.download_fields()
        rows1 = sim.column_rows()
        # a new upload sizes its columns for its own particles: the
        # regrowth of the bunched run does not carry over
        sim.upload_particles(*[st.p[m] for m in states.MEMBERS], st.cell)
        assert sim.column_rows() == rows0
    print(f"column rows {rows0} at upload, {rows1} after the run")
    assert rows1 > rows0
    h, _ = st.run(states.consts_vector(k), STEPS)
    top = np.bincount(got["cell"]).max()
    print(f"{np.dtype(dtype).name}: largest cell count {counts0} at upload, {top} after "
          f"{STEPS} steps")
    assert top > 64, top  # past the columns sized at upload
    assert not np.any(f)
    g = states.canonical(got)
    o = states.canonical(states.oracle_particles(st))
    for m in states.MEMBERS + ("cell",):
        assert np.array_equal(g[m], o[m]), m


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_reserve_rows_keeps_every_particle(dtype):
    """cpic_reserve_column_rows re-lays the columns out on the device: after
    three in-place pushes (columns with holes), the grown columns hold the
    same particles bit for bit (matched by upload index), a request for no
    more rows changes nothing, and the next push equals the oracle's push of
    the state before the regrowth."""
    from cabanapic_amd import decks
    from test_gpu_configs import _assert_same_by_tag, load, oracle_push
    deck = decks.weibel_3d(dtype, n=32, nppc=32)
    sp = deck.step_params(energy_interval=1)
    with load(deck) as sim:
        sim.step(sp, 3)
        f0 = sim.download_fields()
        p0 = sim.download_particles()
        t0 = sim.download_particle_tags()
        d0 = sim.column_rows()
        sim.reserve_column_rows(d0 + 1)
        d1 = sim.column_rows()
        assert d1 >= d0 + 1 and d1 % 64 == 0
        _assert_same_by_tag(sim.download_particles(), sim.download_particle_tags(), p0, t0,
                            "after the regrowth")
        sim.reserve_column_rows(d1)
        assert sim.column_rows() == d1
        sim.step(sp, 1)
        got = sim.download_particles()
        tags = sim.download_particle_tags()
        assert sim.device_errors() == 0
    _assert_same_by_tag(got, tags, oracle_push(deck, f0, p0), t0, "push after the regrowth")
    print(f"{np.dtype(dtype).name}: columns of {d0} rows grown to {d1}")


@pytest.mark.timeout(300)
def test_regrowth_time_full_size():
    """The regrowth on 128^3 cells x 32 particles per cell (67.1M particles,
    FP32) to 1.5x the rows takes less than 50 ms, keeps every particle, and
    the steps after it run."""
    from cabanapic_amd import decks
    from test_gpu_configs import load
    deck = decks.weibel_3d(np.float32)
    sp = deck.step_params(energy_interval=1)
    with load(deck) as sim:
        sim.step(sp, 2)
        sim.synchronize()
        d0, n0 = sim.column_rows(), sim.num_particles()
        t = time.perf_counter()
        sim.reserve_column_rows(d0 * 3 // 2)
        dt = time.perf_counter() - t
        d1 = sim.column_rows()
        sim.step(sp, 2)
        assert sim.device_errors() == 0 and sim.num_particles() == n0
    print(f"regrowth of {n0} particles from {d0} to {d1} rows: {1e3 * dt:.1f} ms")
    assert d1 >= d0 * 3 // 2
    assert dt < 0.05
