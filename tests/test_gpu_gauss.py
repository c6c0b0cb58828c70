"""Gauss's law check on the device (cpic_gauss_residual), FP64.

The residual r = div E - rho / eps0 at every node is checked against a
numpy evaluation of the same definition on the state downloaded from the
device (cloud-in-cell node charge of the particles, backward-difference
divergence of the edge fields, periodic images); the oracle's own test
(tests/test_oracle_physics.py::test_gauss_residual_invariant_over_full_steps)
pins that the charge-conserving step keeps this residual constant, and here
the device's fused step keeps it constant too: in FP64 to rounding, in FP32
within the float rounding of the positions and fields.
"""
import numpy as np
import pytest

import states
from cabanapic_amd import Simulation, CpicError

pytestmark = pytest.mark.gpu

F64 = np.float64


def _name(dtype):
    return "f32" if dtype == np.float32 else "f64"


def numpy_residual(sim, dx, dy, dz, eps0, qsp):
    nx, ny, nz = sim.nx, sim.ny, sim.nz
    f = sim.download_fields().astype(F64).reshape(9, nz + 2, ny + 2, nx + 2)
    p = sim.download_particles()
    X, Y = nx + 2, ny + 2
    ix = p["cell"] % X
    iy = (p["cell"] // X) % Y
    iz = p["cell"] // (X * Y)
    q = p["w"].astype(F64) * F64(qsp)
    Q = np.zeros((nz, ny, nx))
    for a in (0, 1):
        wx = (1 + (2 * a - 1) * p["x"].astype(F64)) / 2
        for b in (0, 1):
            wy = (1 + (2 * b - 1) * p["y"].astype(F64)) / 2
            for c in (0, 1):
                wz = (1 + (2 * c - 1) * p["z"].astype(F64)) / 2
                np.add.at(Q, ((iz - 1 + c) % nz, (iy - 1 + b) % ny, (ix - 1 + a) % nx),
                          q * wx * wy * wz)
    e = [f[m, 1:-1, 1:-1, 1:-1] for m in range(3)]
    div = ((e[0] - np.roll(e[0], 1, axis=2)) / dx + (e[1] - np.roll(e[1], 1, axis=1)) / dy
           + (e[2] - np.roll(e[2], 1, axis=0)) / dz)
    rho = Q / (dx * dy * dz * eps0)
    return div - rho, rho


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=_name)
@pytest.mark.parametrize("grid", [(6, 5, 4, 8), (1, 7, 1, 16), (9, 1, 3, 5)])
def test_gauss_residual_matches_numpy(dtype, grid):
    st, _ = states.random_state(*grid, dtype, seed=31)
    k = states.step_consts(dtype)
    dx, dy, dz = (float(k[n]) for n in ("dx", "dy", "dz"))
    eps0 = float(k["dt"]) / float(k["dt_eps0"])
    with Simulation(st.nx, st.ny, st.nz, st.np, dtype) as sim:
        sim.upload_fields(st.fields)
        sim.upload_particles(*[st.p[m] for m in states.MEMBERS], st.cell)
        m0, mrho, r = sim.gauss_residual(dx, dy, dz, eps0, -1.0, baseline=True, residual=True)
        ref, rho = numpy_residual(sim, dx, dy, dz, eps0, -1.0)
    scale = np.abs(ref).max()
    np.testing.assert_allclose(r, ref, rtol=0, atol=1e-13 * scale)
    assert abs(m0 - np.abs(ref).max()) <= 1e-13 * scale
    assert abs(mrho - np.abs(rho).max()) <= 1e-13 * np.abs(rho).max()


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=_name)
def test_gauss_residual_constant_over_fused_steps(dtype):
    """25 fused steps (load_interpolator, push with move_p, unload,
    advance_b/2, advance_e, advance_b/2) on a random state with fast
    particles: the device's residual changes by at most rounding (FP64:
    1e-11 of the divergence scale, FP32: 1e-4), while particles cross cells
    and E changes by far more."""
    st, _ = states.random_state(6, 5, 4, 8, dtype, seed=9, umax=1.0)
    st.fields[:] = 0
    k = states.step_consts(dtype)
    dx, dy, dz = (float(k[n]) for n in ("dx", "dy", "dz"))
    eps0 = float(k["dt"]) / float(k["dt_eps0"])
    with Simulation(st.nx, st.ny, st.nz, st.np, dtype) as sim:
        sim.upload_fields(st.fields)
        sim.upload_particles(*[st.p[m] for m in states.MEMBERS], st.cell)
        base, rho0 = sim.gauss_residual(dx, dy, dz, eps0, float(k["qsp"]), baseline=True)
        sim.step(dict(k, energy_interval=1), 25)
        change, rho1 = sim.gauss_residual(dx, dy, dz, eps0, float(k["qsp"]), baseline=False)
        _, _, r = sim.gauss_residual(dx, dy, dz, eps0, float(k["qsp"]), baseline=False,
                                     residual=True)
        f = sim.download_fields().astype(F64).reshape(9, st.nz + 2, st.ny + 2, st.nx + 2)
    # E went from zero to a field whose divergence is the charge change
    ediv = np.abs(r).max()
    print(f"{_name(dtype)}: Gauss residual change {change:.3e}, max |rho/eps0| {rho1:.3e}, "
          f"max |E| {np.abs(f[:3]).max():.3e}")
    assert np.abs(f[:3]).max() > 0.1
    tol = 1e-11 if dtype == np.float64 else 1e-4
    assert change <= tol * max(ediv, rho1), (change, ediv, rho1)


def test_gauss_change_needs_a_baseline():
    st, _ = states.random_state(3, 3, 3, 2, np.float64, seed=2)
    with Simulation(st.nx, st.ny, st.nz, st.np, np.float64) as sim:
        sim.upload_particles(*[st.p[m] for m in states.MEMBERS], st.cell)
        with pytest.raises(CpicError):
            sim.gauss_residual(0.1, 0.1, 0.1, 1.0, -1.0, baseline=False)
        with pytest.raises(CpicError):
            sim.gauss_residual(0.0, 0.1, 0.1, 1.0, -1.0, baseline=True)
