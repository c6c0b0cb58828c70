"""Dump snapshots (cpic_snapshot_begin / _wait) and the queued step
(cpic_step_async).

The snapshot holds, in upload order, each particle's in-cell offset x,
momentum ux and voxel: equal to the downloaded particles put in tag order;
its ex line equals ex of the downloaded fields at voxels (1..nx, 1, 1) and
its energies equal cpic_energies.  Steps queued with cpic_step_async equal
cpic_step's, with the two slots used in turn while steps are queued behind
them; a slot begun twice without a wait, or waited for without a begin, is
an error.  A snapshot keeps the particle count of its begin when particles
of another count are uploaded before its wait.
"""
import numpy as np
import pytest

import states
from cabanapic_amd import CpicError
from test_gpu_parity import assert_bitwise, sim_from_state

pytestmark = pytest.mark.gpu


def _check(sim, snap, st):
    p = sim.download_particles()
    t = sim.download_particle_tags()
    order = np.argsort(t)
    assert np.array_equal(t[order], np.arange(t.size))
    assert_bitwise(snap["x"], p["x"][order], "x")
    assert_bitwise(snap["ux"], p["ux"][order], "ux")
    assert_bitwise(snap["cell"], p["cell"][order], "cell")
    f = sim.download_fields()
    X, Y = st.nx + 2, st.ny + 2
    ex = f[0][np.arange(1, st.nx + 1) + X + X * Y]
    assert_bitwise(snap["ex_line"], ex, "ex line")
    e, b = sim.energies()
    assert snap["e_energy"] == e and snap["b_energy"] == b


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_snapshots_follow_queued_steps(dtype):
    st, _ = states.random_state(12, 5, 4, 20, dtype, seed=91)
    sp = dict(states.step_consts(dtype), energy_interval=1)
    with sim_from_state(st) as a, sim_from_state(st) as b:
        a.snapshot_begin(0)
        s0 = a.snapshot_wait(0)
        # before any step: the uploaded particles in their input order
        assert_bitwise(s0["x"], st.p["x"], "x at upload")
        assert_bitwise(s0["ux"], st.p["ux"], "ux at upload")
        assert_bitwise(s0["cell"], st.cell, "cell at upload")
        for step in range(1, 7):
            a.step_async(sp, 1)
            a.snapshot_begin(step % 2)
            if step >= 2:
                snap = a.snapshot_wait((step - 1) % 2)
                # step - 1 on the synchronous context, compared once the
                # queued one has been waited for
                b.step(sp, 1)
                _check(b, snap, st)
        b.step(sp, 1)
        _check(b, a.snapshot_wait(0), st)
        a.synchronize()
        assert a.device_errors() == 0
        with pytest.raises(CpicError):
            a.snapshot_wait(1)
        a.snapshot_begin(1)
        with pytest.raises(CpicError):
            a.snapshot_begin(1)
        a.snapshot_wait(1)


def test_snapshot_keeps_its_layout_across_an_upload():
    # a snapshot begun before an upload of another particle count holds the
    # particles of its begin; the other slot then follows the new count
    st, _ = states.random_state(6, 5, 4, 20, np.float32, seed=93)
    st2, _ = states.random_state(6, 5, 4, 7, np.float32, seed=94)
    assert st2.np < st.np
    with sim_from_state(st) as sim:
        sim.snapshot_begin(0)
        sim.upload_particles(*[st2.p[m] for m in ("x", "y", "z", "ux", "uy", "uz", "w")],
                             st2.cell)
        sim.snapshot_begin(1)
        for slot, s in ((0, st), (1, st2)):
            snap = sim.snapshot_wait(slot)
            assert snap["x"].size == s.np
            assert_bitwise(snap["x"], s.p["x"], f"x, slot {slot}")
            assert_bitwise(snap["ux"], s.p["ux"], f"ux, slot {slot}")
            assert_bitwise(snap["cell"], s.cell, f"cell, slot {slot}")
        assert sim.device_errors() == 0
