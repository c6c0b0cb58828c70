"""Failure handling of the push and particle identity through the columns.

* A cell that receives more particles than its column holds, the last of
  them through the spill fallback (records beyond a (cell, face) list's
  8 slots): the push raises the column-full flag and leaves every column
  count within the capacity; the next push is refused with CPIC_ERR_DEVICE;
  a download reports the lost particles as an error (no fault, no
  out-of-bounds access); a new upload clears the state and pushes again.
* Tags: every particle keeps its upload index through packing and in-place
  pushes, crossings, records and spills, so the downloaded particles can be
  matched one for one with the oracle's, which keeps the upload order.
"""
import numpy as np
import pytest

import oracle_lib
import states
from cabanapic_amd import CpicError, Simulation

pytestmark = pytest.mark.gpu

N = 16  # cells per axis
K = dict(qdt_2mc=-0.05, cdt_dx=0.45, cdt_dy=0.4, cdt_dz=0.35, qsp=-1.0)


def _vox(ix, iy, iz):
    return ix + (N + 2) * (iy + (N + 2) * iz)


def crowding_state(dtype):
    """Cell T = (8, 8, 8) holds 16 particles at rest and receives 16 across
    each x face (one crossing) and 16 from each of two x-y edges (two
    crossings: +x, then +y or -y).  The 16 second crossings of each edge
    share one (cell, face) record list of 8 slots, so 8 of each take the
    spill fallback.  Arrive and finish passes fill T to exactly 64, the
    column capacity (16 per cell at upload); the spills overflow it.  Four
    particles at rest in every other cell make the context large enough for
    the usual capacity rule (2.5x the largest cell, at least +48 rows)."""
    parts = []

    def add(cell, n, x, y, z, ux, uy, uz):
        parts.extend([(cell, x, y, z, ux, uy, uz)] * n)

    t = _vox(8, 8, 8)
    add(t, 16, 0, 0, 0, 0, 0, 0)
    add(_vox(7, 8, 8), 16, 0.9, 0, 0, 0.9, 0, 0)
    add(_vox(9, 8, 8), 16, -0.9, 0, 0, -0.9, 0, 0)
    add(_vox(7, 7, 8), 16, 0.9, 0.9, 0, 0.9, 0.9, 0)
    add(_vox(7, 9, 8), 16, 0.9, -0.9, 0, 0.9, -0.9, 0)
    used = {t, _vox(7, 8, 8), _vox(9, 8, 8), _vox(7, 7, 8), _vox(7, 9, 8)}
    for iz in range(1, N + 1):
        for iy in range(1, N + 1):
            for ix in range(1, N + 1):
                v = _vox(ix, iy, iz)
                if v not in used:
                    add(v, 4, -0.5, 0.25, 0.5, 0, 0, 0)
    a = np.array(parts)
    st = oracle_lib.OracleState(N, N, N, len(parts), dtype)
    st.cell[:] = a[:, 0].astype(np.int32)
    for i, m in enumerate(("x", "y", "z", "ux", "uy", "uz")):
        st.p[m][:] = a[:, 1 + i].astype(dtype)
    st.p["w"][:] = dtype(1)
    return st, t


def _upload(sim, st):
    sim.upload_particles(*[st.p[m] for m in states.MEMBERS], st.cell)


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_spill_into_full_column_fails_cleanly(dtype):
    st, t = crowding_state(dtype)
    # the oracle puts 80 particles into T
    ref = st.copy()
    ref.load_interpolator()
    ref.clear_accumulator()
    ref.push(*(dtype(K[n]) for n in ("qdt_2mc", "cdt_dx", "cdt_dy", "cdt_dz", "qsp")))
    assert np.count_nonzero(ref.cell == t) == 80
    with Simulation(N, N, N, st.np, dtype) as sim:
        _upload(sim, st)
        assert sim.column_rows() == 64
        sim.load_interpolator_array()
        sim.clear_accumulator_array()
        sim.push(**K)  # asynchronous: the failure is seen afterwards
        flags = sim.device_errors()
        assert flags & 4, flags  # column full
        with pytest.raises(CpicError) as e:
            sim.push(**K)
        assert e.value.code == 4  # CPIC_ERR_DEVICE
        with pytest.raises(CpicError) as e:
            sim.download_particles()  # 16 particles could not be stored
        assert e.value.code == 4
        # a fresh upload clears the flags; the same state fails the same way
        _upload(sim, st)
        assert sim.device_errors() == 0
        sim.load_interpolator_array()
        sim.clear_accumulator_array()
        sim.push(**K)
        assert sim.device_errors() & 4
    # with room for the crowd the push matches the oracle
    with Simulation(N, N, N, st.np, dtype, column_rows=128) as sim:
        _upload(sim, st)
        assert sim.column_rows() == 128
        sim.load_interpolator_array()
        sim.clear_accumulator_array()
        sim.push(**K)
        assert sim.device_errors() == 0
        got = sim.download_particles()
    assert np.count_nonzero(got["cell"] == t) == 80
    g = states.canonical(got)
    o = states.canonical(states.oracle_particles(ref))
    for m in states.MEMBERS + ("cell",):
        assert np.array_equal(g[m], o[m]), m


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_tags_follow_particles(dtype):
    """Four pushes of a random state on its interpolators (a packing push,
    then in-place ones; crossings of up to three faces): after each, every
    downloaded particle equals, bit for bit, the oracle's particle of the
    same upload index."""
    st, k = states.random_state(7, 6, 5, 24, dtype, seed=41, umax=1.0)
    st.load_interpolator()
    st.clear_accumulator()
    with Simulation(st.nx, st.ny, st.nz, st.np, dtype) as sim:
        sim.upload_fields(st.fields)
        _upload(sim, st)
        sim.load_interpolator_array()
        sim.clear_accumulator_array()
        for _ in range(4):
            sim.push(**k)
            st.push(**k)
            got = sim.download_particles()
            tags = sim.download_particle_tags()
            assert np.array_equal(np.sort(tags), np.arange(st.np))
            for m in states.MEMBERS:
                a = np.ascontiguousarray(got[m])
                b = np.ascontiguousarray(st.p[m][tags])
                assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), m
            assert np.array_equal(got["cell"], st.cell[tags])
        assert sim.device_errors() == 0
