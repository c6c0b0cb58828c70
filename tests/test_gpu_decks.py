"""The reference's own tests, run through the C++ driver on the GPU.

* energy regression (tests/energy_comparison): the two-stream EM deck for
  6000 steps, energy history checked by the deck's finalizer against the
  gold file of the same precision with the reference's 10 % rule (the
  executable exits 1 on failure); in double precision the history must also
  match the gold file to its printed digits over the whole run.  The float
  E rule is order sensitive (see FLOAT_E_ENVELOPE).
* deck-override smoke test (tests/decks): a deck with its own particle
  initializer runs 30 steps and exits 0.
* the reference's own deck sources, compiled unchanged against the deck
  interface where the reference tree is present (the executables travel
  with the tree): its regression test tests/energy_comparison/2stream-em.cxx
  with its own gold check, and decks/2stream-short.cxx, custom_init.cxx and
  dioctron_3d.cxx, which run to the end with finite energies.
* the electrostatic build (the reference's SOLVER_TYPE=ES): the 1D deck's
  energies.txt holds step, time and E, and E follows the oracle's ES stage
  loop from the deck's own initial state.
"""
import os
import subprocess

import numpy as np
import pytest

import energy_rule
import host_driver

pytestmark = pytest.mark.gpu


def run(deck, real, cwd, *args, timeout=300):
    exe = host_driver.exe(deck, real)
    print(f"running {os.path.basename(exe)} {' '.join(args)}", flush=True)
    r = subprocess.run([exe, *args], cwd=cwd, capture_output=True, text=True, timeout=timeout)
    return r


def _rule_errors(stdout):
    """The two "Max found err was X%" figures the deck's finalizer prints
    (E first, then B), as fractions."""
    errs = [float(l.split("was ")[1].split("%")[0]) / 100 for l in stdout.splitlines()
            if l.startswith("Max found err was")]
    assert len(errs) == 2, stdout[-2000:]
    return errs


# In float the E window of the rule lies in the nonlinear phase, where the
# history is chaotic in the summation order of the deposit: the float oracle
# itself fails the 10 % E rule for 5 of 23 random particle orders (max errors
# 3-16 %, tests/test_oracle_gold.py::test_float_e_rule_is_order_sensitive)
# while B stays below 1 %.  The float run must therefore pass the B rule and
# keep the E error inside that order-only envelope.
FLOAT_E_ENVELOPE = 0.20


@pytest.mark.parametrize("real", ["float", "double"])
def test_energy_regression_deck(real, tmp_path):
    r = run("two_stream_em_regression", real, tmp_path)
    err_e, err_b = _rule_errors(r.stdout)
    print(f"rule errors E {err_e:.4f} B {err_b:.4f} rc {r.returncode}", flush=True)
    assert "B Test Pass: 1" in r.stdout
    if real == "double":
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        assert "E Test Pass: 1" in r.stdout
    else:
        assert r.returncode in (0, 1), r.stderr[-2000:]
        assert err_e < FLOAT_E_ENVELOPE and err_b < 0.02, (err_e, err_b)
    h = np.loadtxt(tmp_path / "energies.txt")
    g = energy_rule.gold(real)
    assert h.shape == g.shape
    if real == "double":
        rel = np.abs(h[:, 2:4] - g[:, 2:4]) / g[:, 2:4]
        assert rel.max() < 1e-5, rel.max()
    else:
        rel = np.abs(h[:10, 2:4] - g[:10, 2:4]) / g[:10, 2:4]
        assert rel.max() < 5e-6, rel.max()
    rate = energy_rule.growth_rate(h[:, 3], h[1, 1] - h[0, 1])
    assert abs(rate - 0.2418) < 0.005, rate


def test_energy_regression_deck_stage_by_stage(tmp_path):
    """The same deck driven one C-ABI call per stage (the reference's
    sequence of calls) instead of the fused step."""
    r = run("two_stream_em_regression", "float", tmp_path, "--stages")
    err_e, err_b = _rule_errors(r.stdout)
    assert r.returncode in (0, 1), r.stderr[-2000:]
    assert "B Test Pass: 1" in r.stdout and err_e < FLOAT_E_ENVELOPE, (err_e, err_b)


@pytest.mark.parametrize("real", ["float", "double"])
def test_custom_initializer_deck(real, tmp_path):
    r = run("beam_1d_custom_init", real, tmp_path)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "Using a custom particle initializer" in r.stdout
    h = np.loadtxt(tmp_path / "energies.txt")
    assert h.shape == (30, 4) and np.all(np.isfinite(h))


@pytest.mark.parametrize("real", ["float", "double"])
def test_es_solver_deck(real, tmp_path):
    r = run("beam_1d_custom_init_es", real, tmp_path)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    h = np.loadtxt(tmp_path / "energies.txt")
    assert h.shape == (30, 3) and np.all(np.isfinite(h))
    import oracle_lib
    k, p = host_driver.dump("beam_1d_custom_init_es", real, str(tmp_path / "init.bin"))
    dt = np.float32 if real == "float" else np.float64
    st = oracle_lib.OracleState(32, 1, 1, p["x"].size, dt)
    for m in ("x", "y", "z", "ux", "uy", "uz", "w"):
        st.p[m][:] = p[m]
    st.cell[:] = p["cell"]
    R = lambda n: dt(k[n])
    e = []
    for _ in range(30):
        st.load_interpolator()
        st.clear_accumulator()
        st.push(R("qdt_2mc"), R("cdt_dx"), R("cdt_dy"), R("cdt_dz"), R("qsp"))
        st.unload(R("dx"), R("dy"), R("dz"), R("dt"))
        st.es_advance_e(R("dt_eps0"))
        e.append(st.es_energy()[0])
    e = np.array(e)
    # energies.txt prints 6 significant digits.  In float the history is
    # sensitive to the order of the current sums alone: the oracle's own run
    # with the particle list permuted (6 random orders) moves by up to
    # 2.1e-3 of the largest energy over the 30 steps (1.3e-12 in double).
    rel = np.abs(h[:, 2] - e) / np.abs(e).max()
    assert rel.max() < (5e-3 if real == "float" else 2e-5), rel.max()


def _ref_exe(name, real):
    exe = os.path.join(host_driver.HOST, "build", f"cbnpic_{name}_{real}")
    if not os.path.exists(exe):
        pytest.skip("the reference's decks are built only where the reference tree is present")
    return exe


@pytest.mark.parametrize("real", ["float", "double"])
def test_reference_regression_test_unchanged(real, tmp_path):
    """The reference's regression test source, built unchanged: its own
    finalizer checks the 6000-step history against the gold file with its
    10 % rule.  Double passes both rules and matches the gold to its printed
    digits; float passes B and keeps E inside the order-only envelope."""
    exe = _ref_exe("ref_2stream_em", real)
    print(f"running {os.path.basename(exe)}", flush=True)
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    err_e, err_b = _rule_errors(r.stdout)
    assert "B Test Pass: 1" in r.stdout
    h = np.loadtxt(tmp_path / "energies.txt")
    g = energy_rule.gold(real)
    assert h.shape == g.shape
    if real == "double":
        assert r.returncode == 0 and "E Test Pass: 1" in r.stdout, r.stdout[-2000:]
        assert (np.abs(h[:, 2:4] - g[:, 2:4]) / g[:, 2:4]).max() < 1e-5
    else:
        assert r.returncode in (0, 1) and err_e < FLOAT_E_ENVELOPE, (err_e, err_b)


@pytest.mark.parametrize("name,steps", [("ref_custom_init", 30), ("ref_dioctron_3d", 20000)])
def test_reference_decks_run_unchanged(name, steps, tmp_path):
    exe = _ref_exe(name, "double")
    print(f"running {os.path.basename(exe)}", flush=True)
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    h = np.loadtxt(tmp_path / "energies.txt")
    assert h.shape == (steps, 4) and np.all(np.isfinite(h))


def test_reference_deck_with_ghost_particles_is_rejected(tmp_path):
    """decks/2stream-short.cxx is 1D in x (32 x 1 x 1 cells) but keeps the
    default initializer, which spreads the beams along y
    (src/input/deck.h:79-158): particle 100 onwards land in ghost voxels
    (y = 2 of a one-cell-high grid).  The reference pushes them on the zero
    ghost interpolators; this build refuses particles outside the interior
    at upload, with an error naming the first one, instead of running."""
    exe = _ref_exe("ref_2stream_short", "double")
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 1
    assert "particle 100 has non-interior cell" in r.stderr, r.stderr[-2000:]


def _expected_particle_line(p, dx, nx, ny, real):
    """One partloc line in the reference's format (src/helpers.h:31-70),
    computed with its arithmetic: x = xmin + (ix - 1 + (px + 1.0) * 0.5) * dx
    in double from real_t inputs, rounded to real_t."""
    R = np.float32 if real == "float" else np.float64
    X, Y = nx + 2, ny + 2
    ix = p["cell"] % X
    x = (ix - 1 + (p["x"].astype(np.float64) + 1.0) * 0.5) * np.float64(R(dx))
    x = x.astype(R)
    return "".join(f"{float(a):e}  {float(b):e} " for a, b in zip(x, p["ux"]))


def _partloc(path):
    """{step: values} of a partloc file (time first, then x, ux pairs)."""
    lines = path.read_text().split("\n")
    out = {}
    for i in range(0, len(lines) - 1, 2):
        if not lines[i].startswith("#step="):
            break
        out[int(lines[i][6:])] = np.array(lines[i + 1].split(), float)
    return lines, out


@pytest.mark.parametrize("real", ["float", "double"])
def test_reference_dumps(real, tmp_path):
    """--dumps writes the reference's partloc and ex1d files
    (example/example.cpp:59-60,130-132,273-277): a step-0 particle line
    equal to the reference's formula applied to the deck's initial state,
    then one block per step with every particle and the nx ex samples.
    The dumps run queues the steps without waiting and formats step k from a
    snapshot while later steps run: its energies.txt equals the plain fused
    run's, its partloc and ex1d equal those of the stage-by-stage run (which
    takes each snapshot synchronously), and the particles keep the order of
    the initial state (every particle moves by at most c dt per step)."""
    deck = "beam_1d_custom_init"
    r = run(deck, real, tmp_path, "--dumps")
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    k, p = host_driver.dump(deck, real, str(tmp_path / "init.bin"))
    n = p["cell"].size
    lines, dumps = _partloc(tmp_path / "partloc")
    assert lines[0] == "#step=0"
    assert lines[1] == "0 " + _expected_particle_line(p, k["dx"], 32, 1, real)
    assert sorted(dumps) == list(range(31))
    for step in range(1, 31):
        vals = dumps[step]
        assert len(vals) == 1 + 2 * n and float(vals[0]) == pytest.approx(step * k["dt"], rel=1e-6)
        assert np.all(np.isfinite(vals))
    # particle order: x of particle i moves by at most c dt (c = 1) per step,
    # across the periodic wrap of the 32-cell box
    length = 32 * k["dx"]
    for step in range(1, 31):
        d = np.abs(dumps[step][1::2] - dumps[step - 1][1::2])
        d = np.minimum(d, length - d)
        assert d.max() <= k["dt"] * (1 + 1e-5), (step, d.max())
    blocks = (tmp_path / "ex1d").read_text().split("#step=")[1:]
    assert len(blocks) == 30
    for step, b in enumerate(blocks, 1):
        rows = [l.split() for l in b.strip().split("\n")[1:]]
        assert b.startswith(f"{step}\n") and len(rows) == 32
        xs = np.array([float(r_[0]) for r_ in rows])
        assert np.allclose(xs, (np.arange(1, 33) - 0.5) * k["dx"], rtol=1e-6)

    plain = tmp_path / "plain"
    plain.mkdir()
    r = run(deck, real, plain)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert (plain / "energies.txt").read_text() == (tmp_path / "energies.txt").read_text()
    staged = tmp_path / "staged"
    staged.mkdir()
    r = run(deck, real, staged, "--stages", "--dumps")
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    _, sd = _partloc(staged / "partloc")
    assert sorted(sd) == sorted(dumps)
    for step in dumps:
        assert np.array_equal(sd[step], dumps[step]), step
    ex_a = np.array((tmp_path / "ex1d").read_text().replace("#step=", " ").split(), float)
    ex_b = np.array((staged / "ex1d").read_text().replace("#step=", " ").split(), float)
    assert np.array_equal(ex_a, ex_b)


def test_column_deck_host_driver_matches_python_deck(tmp_path):
    """The electron-column deck (192 x 192 x 64 cells, 32 per cell, 1000
    steps) through the C++ driver, whose Plasma_Initializer and
    Column_Field_Initializer call the library's loaders: it exits 0, its
    Gauss residual drifts by less than 1e-4 of max rho / eps0 over the run,
    and its energies.txt (6 significant digits) equals the energy history
    of the Python deck's fused run on the same library."""
    from cabanapic_amd import Simulation, decks
    r = run("dioctron_3d", "float", tmp_path, "--chunk", "1000", timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    gauss = [l for l in r.stderr.splitlines() if l.startswith("#Gauss residual")]
    assert len(gauss) == 1, r.stderr[-2000:]
    rel = float(gauss[0].split("relative ")[1].rstrip(")"))
    print(gauss[0])
    assert rel < 1e-4
    h = np.loadtxt(tmp_path / "energies.txt")
    d = decks.dioctron_3d(np.float32)
    assert h.shape == (d.num_steps, 4) and np.all(np.isfinite(h))
    parts = d.load_particles()
    with Simulation(d.nx, d.ny, d.nz, d.num_particles, d.R) as sim:
        sim.upload_particles(*[parts[m] for m in ("x", "y", "z", "ux", "uy", "uz", "w")],
                             parts["cell"])
        del parts
        sim.upload_fields(d.load_fields())
        py = sim.step(d.step_params(energy_interval=1), d.num_steps)
        assert sim.device_errors() == 0
    for col, what in ((2, "E"), (3, "B")):
        ref = py[:, col - 2].astype(np.float32).astype(np.float64)
        err = np.abs(h[:, col] - ref) / np.abs(ref)
        print(f"{what}: max rel. difference to the Python deck {err.max():.2e}")
        assert err.max() < 1e-5, what
