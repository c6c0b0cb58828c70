"""The reverse-mode adjoint (cpic_forward / cpic_backward) on the GPU.

* Against the oracle's FP64 reference adjoint (oracle/oracle_adjoint.inc,
  pinned by central differences and dot-product identities): one step of a
  random state on a 4^3 and a 5 x 4 x 6 grid, random seeds on every field
  and particle component, FP64 and FP32 primal; the GPU and the oracle start
  from the same state (the push is bit-identical), so the FP64 adjoints
  agree to their summation order.  Five steps in FP64 with checkpoints
  every 2 steps, against the oracle's chain of one-step adjoints.
* Checkpoint interval: the gradient with checkpoints every 1, 3 and 10 of
  10 steps is the same.
* Dot-product identity with the tangent-linear step over 10 FP64 steps:
  the tangent's derivative of the final E energy along a random direction v
  (particles and fields) equals <v, adjoint of the initial state> for the
  seed dW_E/dE = E, to 1e-10.
* The differentiation configuration (3D two-stream 32^3, 2 x 16 ppc,
  v0 = 0.2, amplitude 1e-3, 500 steps, checkpoints every 50): the adjoint's
  d(E energy)/d(v0) and d/d(amplitude) equal the tangent's to 1e-8 in
  FP64; the FP32-primal gradients are printed.
"""
import numpy as np
import pytest

import oracle_lib
import states
from cabanapic_amd import Adjoint, Simulation, TangentSimulation, decks

pytestmark = pytest.mark.gpu

F32, F64 = np.float32, np.float64
MEMBERS = states.MEMBERS


def _name(dtype):
    return "f32" if dtype == F32 else "f64"


def _state(grid, dtype, seed):
    st, _ = states.random_state(*grid, 8, dtype, seed=seed, umax=0.8)
    st.p["w"][:] = (st.p["w"] * 0.05).astype(dtype)
    k = states.step_consts(dtype)
    return st, k


def _sim(st):
    sim = Simulation(st.nx, st.ny, st.nz, st.np, st.dtype)
    sim.upload_fields(st.fields)
    sim.upload_particles(*[st.p[m] for m in MEMBERS], st.cell)
    return sim


def _close(a, b, tol):
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b).max() / scale
    return err <= tol, err


@pytest.mark.parametrize("dtype", [F64, F32], ids=_name)
@pytest.mark.parametrize("grid", [(4, 4, 4), (5, 4, 6)], ids=lambda g: "x".join(map(str, g)))
def test_one_step_matches_oracle_adjoint(dtype, grid):
    st, k = _state(grid, dtype, seed=51)
    rng = np.random.default_rng(52)
    seed_f = rng.uniform(-1, 1, (9, st.ng))
    seed_p = rng.uniform(-1, 1, (6, st.np))
    with _sim(st) as sim, Adjoint(sim) as adj:
        adj.forward(dict(k, energy_interval=0), 1, 1, history=False)
        gf, gp = adj.backward(seed_f, seed_p)
        assert sim.device_errors() == 0
    lf = seed_f.copy()
    lp = st.step_adjoint(states.consts_vector(k), lf, seed_p)
    ok_f, ef = _close(gf, lf, 1e-12)
    ok_p, ep = _close(gp, lp, 1e-12)
    print(f"{_name(dtype)} {grid}: fields {ef:.2e}, particles {ep:.2e}")
    assert ok_f and ok_p, (ef, ep)


def test_five_steps_match_oracle_chain():
    st, k = _state((5, 4, 6), F64, seed=61)
    consts = states.consts_vector(k)
    rng = np.random.default_rng(62)
    seed_f = rng.uniform(-1, 1, (9, st.ng))
    seed_p = rng.uniform(-1, 1, (6, st.np))
    with _sim(st) as sim, Adjoint(sim) as adj:
        adj.forward(dict(k, energy_interval=0), 5, 2, history=False)
        gf, gp = adj.backward(seed_f, seed_p)
        assert sim.device_errors() == 0
    traj = [st.copy()]
    for _ in range(5):
        s = traj[-1].copy()
        s.run(consts, 1, hist=False)
        traj.append(s)
    lf = seed_f.copy()
    lp = seed_p
    for s in reversed(traj[:-1]):
        lp = s.step_adjoint(consts, lf, lp)
    ok_f, ef = _close(gf, lf, 1e-9)
    ok_p, ep = _close(gp, lp, 1e-9)
    print(f"five steps: fields {ef:.2e}, particles {ep:.2e}")
    assert ok_f and ok_p, (ef, ep)


def _e_energy_seed(sim):
    """dW_E/dfields for W_E = 0.5 sum over interior cells of E^2."""
    f = sim.download_fields().astype(F64)
    seed = np.zeros_like(f)
    m = states.interior_mask(sim.nx, sim.ny, sim.nz)
    seed[:3, m] = f[:3, m]
    return seed


def test_checkpoint_interval_does_not_change_the_gradient():
    st, k = _state((5, 4, 6), F64, seed=71)
    grads = []
    for interval in (1, 3, 10):
        with _sim(st) as sim, Adjoint(sim) as adj:
            adj.forward(dict(k, energy_interval=0), 10, interval, history=False)
            seed = _e_energy_seed(sim)
            grads.append(adj.backward(seed))
            mem = adj.memory()
        print(f"interval {interval}: {mem}")
    for gf, gp in grads[1:]:
        assert _close(gf, grads[0][0], 1e-11)[0]
        assert _close(gp, grads[0][1], 1e-11)[0]


def test_dot_product_identity_with_tangent():
    """<J v, w> = <v, J^T w> over 10 FP64 steps: the tangent of the final
    E energy along v against the adjoint contracted with v."""
    st, k = _state((6, 5, 4), F64, seed=81)
    rng = np.random.default_rng(82)
    vp = rng.uniform(-1, 1, (6, st.np))
    vf = np.zeros((9, st.ng))
    vf[:6] = rng.uniform(-1, 1, (6, st.ng))  # E and B (J is overwritten by the step)
    nsteps = 10
    with TangentSimulation(st.nx, st.ny, st.nz, st.np, F64) as ts:
        ts.primal.upload_fields(st.fields)
        ts.tangent.upload_fields(vf)
        ts.upload_particles(*[st.p[m] for m in MEMBERS], st.cell, vp)
        hist = ts.step(dict(k, energy_interval=1), nsteps)
    tangent = hist[-1, 2]
    with _sim(st) as sim, Adjoint(sim) as adj:
        adj.forward(dict(k, energy_interval=1), nsteps, 4, history=False)
        gf, gp = adj.backward(_e_energy_seed(sim))
    adjoint = float(np.vdot(gp, vp) + np.vdot(gf, vf))
    rel = abs(adjoint - tangent) / abs(tangent)
    print(f"<J v, w> (tangent) {tangent:.15e}, <v, J^T w> (adjoint) {adjoint:.15e}, rel {rel:.2e}")
    assert rel < 1e-10


def _diff_factory(dtype, **kw):
    args = dict(n=32, nppc=32, steps=500, v0=0.2, amp=1e-3, length=2.0 * np.pi * 0.2 / 0.5)
    args.update(kw)
    return decks.two_stream_3d(dtype, **args)


def _loader_tangent(name, value, rel=1e-6):
    """d(ux)/d(parameter) of the FP64 particle loader by central differences
    (ux is linear in the drift and in the amplitude: only rounding enters)."""
    lo = _diff_factory(F64, **{name: value * (1 - rel)}).load_particles()
    hi = _diff_factory(F64, **{name: value * (1 + rel)}).load_particles()
    t = np.zeros((6, lo["cell"].size))
    t[3] = (hi["ux"] - lo["ux"]) / (2 * rel * value)
    return t


@pytest.mark.timeout(900)
def test_differentiation_config_adjoint_equals_tangent():
    params = (("v0", 0.2), ("amp", 1e-3))
    tans = {p: _loader_tangent(p, v) for p, v in params}
    out = {}
    for dtype in (F64, F32):
        deck = _diff_factory(dtype)
        parts = deck.load_particles()
        sp = deck.step_params(energy_interval=1)
        tang = {}
        for p, _ in params:
            with TangentSimulation(deck.nx, deck.ny, deck.nz, deck.num_particles, dtype) as ts:
                ts.upload_particles(*[parts[m] for m in MEMBERS], parts["cell"], tans[p])
                tang[p] = ts.step(sp, deck.num_steps)[-1, 2]
        sim = Simulation(deck.nx, deck.ny, deck.nz, deck.num_particles, dtype)
        sim.upload_particles(*[parts[m] for m in MEMBERS], parts["cell"])
        with sim, Adjoint(sim) as adj:
            h = adj.forward(sp, deck.num_steps, 50)
            _, gp = adj.backward(_e_energy_seed(sim))
            mem = adj.memory()
            assert sim.device_errors() == 0
        adjg = {p: float(np.vdot(gp, tans[p])) for p, _ in params}
        out[dtype] = (tang, adjg)
        print(f"{_name(dtype)}: final E energy {h[-1, 0]:.6e}; checkpoints {mem}")
        for p, _ in params:
            print(f"  dW_E/d{p}: tangent {tang[p]: .12e}, adjoint {adjg[p]: .12e}, "
                  f"rel {abs(adjg[p] - tang[p]) / abs(tang[p]):.2e}")
    tang, adjg = out[F64]
    for p, _ in params:
        assert abs(adjg[p] - tang[p]) <= 1e-8 * abs(tang[p]), p


def _column_run(deck, nsteps):
    """Final fields of nsteps FP64 steps of an electron-column deck."""
    parts = deck.load_particles()
    with Simulation(deck.nx, deck.ny, deck.nz, deck.num_particles, deck.R) as sim:
        sim.upload_particles(*[parts[m] for m in MEMBERS], parts["cell"])
        sim.upload_fields(deck.load_fields())
        sim.step(deck.step_params(energy_interval=0), nsteps, history=False)
        return sim.download_fields()


@pytest.mark.timeout(600)
def test_column_mode_gradient_equals_tangent():
    """A reduced electron column (48 x 48 x 8 cells, 8 ppc, guide field
    B_z = 1, m = 3 perturbation), 60 FP64 steps with checkpoints every 16:
    the adjoint's dA_3/dB_z of the m = 3 amplitude of the radial E field
    equals the tangent-linear step's to 1e-8; central differences of the
    forward run (B_z +- 1e-6 relative) are printed beside them."""
    nsteps = 60
    deck = decks.dioctron_3d(F64, nx=48, ny=48, nz=8, nppc=8, steps=nsteps)
    parts = deck.load_particles()
    fields = deck.load_fields()
    tf, tp = decks.column_guide_field_tangent(deck)
    sp = deck.step_params(energy_interval=1)
    with Simulation(deck.nx, deck.ny, deck.nz, deck.num_particles, F64) as sim, \
            Adjoint(sim) as adj:
        sim.upload_particles(*[parts[m] for m in MEMBERS], parts["cell"])
        sim.upload_fields(fields)
        adj.forward(sp, nsteps, 16)
        A, seed = decks.column_mode_amplitude(deck, sim.download_fields())
        gf, gp = adj.backward(seed)
        assert sim.device_errors() == 0
    g_adj = float(np.vdot(gf, tf) + np.vdot(gp, tp))
    with TangentSimulation(deck.nx, deck.ny, deck.nz, deck.num_particles, F64) as ts:
        ts.primal.upload_fields(fields)
        ts.tangent.upload_fields(tf)
        ts.upload_particles(*[parts[m] for m in MEMBERS], parts["cell"], tp)
        ts.step(sp, nsteps)
        g_tan = float(np.vdot(seed, ts.tangent.download_fields()))
    eps = 1e-6
    amps = []
    for sgn in (1, -1):
        d = decks.dioctron_3d(F64, nx=48, ny=48, nz=8, nppc=8, steps=nsteps, bz=1.0 + sgn * eps)
        amps.append(decks.column_mode_amplitude(d, _column_run(d, nsteps))[0])
    fd = (amps[0] - amps[1]) / (2 * eps)
    rel = abs(g_adj - g_tan) / abs(g_tan)
    print(f"A_3 = {A:.6e}; dA_3/dB_z: adjoint {g_adj:.12e}, tangent {g_tan:.12e} (rel {rel:.2e}), "
          f"central differences {fd:.6e}")
    assert rel < 1e-8
