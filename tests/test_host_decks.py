"""CPU checks of the C++ deck layer (cabanapic_amd/host): the regression
deck's derived constants and default particle initializer reproduce the
oracle's restatement of src/input/deck.h and example/example.cpp bit for bit,
and the Python decks agree with the C++ decks on the derived constants."""
import os
import subprocess

import numpy as np
import pytest

import host_driver
import oracle_lib
from cabanapic_amd import decks


@pytest.mark.parametrize("real", ["float", "double"])
def test_regression_deck_initial_state_matches_oracle(real, tmp_path):
    dtype = np.float32 if real == "float" else np.float64
    k, p = host_driver.dump("two_stream_em_regression", real, str(tmp_path / "init.bin"))
    d, st = oracle_lib.setup_2stream_em(dtype)
    names = ["qdt_2mc", "cdt_dx", "cdt_dy", "cdt_dz", "qsp", "dx", "dy", "dz", "dt", "px", "py",
             "pz", "dt_eps0"]
    for i, n in enumerate(names):
        assert k[n] == d["consts"][i], n
    assert k["dxp"] == d["dxp"] and k["we"] == d["we"] and k["num_steps"] == d["num_steps"]
    assert p["cell"].size == st.np
    ui = np.uint32 if dtype == np.float32 else np.uint64
    for m in ("x", "y", "z", "ux", "uy", "uz", "w"):
        assert np.array_equal(p[m].view(ui), st.p[m].view(ui)), m
    assert np.array_equal(p["cell"], st.cell)


@pytest.mark.parametrize("real", ["float", "double"])
@pytest.mark.parametrize("deck,factory", [("weibel_3d", decks.weibel_3d),
                                          ("two_stream_3d", decks.two_stream_3d),
                                          ("two_stream_1d", decks.two_stream_1d),
                                          ("dioctron_3d", decks.dioctron_3d)])
def test_python_and_cpp_decks_agree(real, deck, factory, tmp_path):
    dtype = np.float32 if real == "float" else np.float64
    k, _ = host_driver.dump(deck, real, str(tmp_path / "c.bin"), particles=False)
    py = factory(dtype)
    pk = py.driver_constants()
    for n in ("qdt_2mc", "cdt_dx", "cdt_dy", "cdt_dz", "qsp", "dx", "dy", "dz", "dt", "px", "py",
              "pz", "dt_eps0", "dxp", "we"):
        assert k[n] == float(pk[n]), (n, k[n], float(pk[n]))


def test_plasma_loader_matches_python_deck_particles(tmp_path):
    """The C++ Plasma_Initializer and the Python deck use the same library
    loader: same particles for a reduced two-stream grid."""
    d = decks.two_stream_3d(np.float32, n=8, nppc=4)
    p = d.load_particles()
    assert p["cell"].size == 8 ** 3 * 4
    # counter-streaming: alternate drift signs, mean |ux| ~ drift
    assert np.mean(p["ux"][0::2]) > 0.15 and np.mean(p["ux"][1::2]) < -0.15
    assert np.all(np.abs(p["x"]) <= 1) and np.all(np.diff(p["cell"]) >= 0)


def test_column_deck_fields_and_drifts():
    """The electron-column loader (a reduced 48 x 48 x 4 grid of the
    dioctron configuration): its radial field is the closed form of
    Gauss's law for the ring, checked against numerical quadrature of
    (1/r) int_0^r rho(s) s ds at the ex and ey Yee edges; particle weights
    follow the ring profile with the cos(3 theta) perturbation; every
    particle drifts at E x B / B^2 of the field at its own position."""
    from scipy.integrate import quad
    d = decks.dioctron_3d(np.float64, nx=48, ny=48, nz=4, nppc=4)
    pl = d.plasma
    r0, sig, rho0, bz = pl["ring_radius"], pl["ring_width"], pl["column_charge"], pl["guide_bz"]
    L, dx = float(d.len_x), float(d.dx)

    def er(r):
        return rho0 / r * quad(lambda s: np.exp(-((s - r0) / sig) ** 2) * s, 0.0, r,
                               epsabs=0, epsrel=1e-13, limit=200)[0]

    f = d.load_fields()
    X, Y = d.nx + 2, d.ny + 2
    N = d.num_cells
    assert f.shape == (9, N)
    # the closed form cancels near the axis, where the field vanishes: the
    # bound is relative to the peak field
    tol = 1e-12 * np.abs(f[:2]).max()
    rng = np.random.default_rng(5)
    for v in rng.integers(0, X * Y, 40):
        ix, iy = v % X, v // X
        xc, xl = (ix - 0.5) * dx - L / 2, (ix - 1.0) * dx - L / 2
        yc, yl = (iy - 0.5) * dx - L / 2, (iy - 1.0) * dx - L / 2
        r = np.hypot(xc, yl)
        assert abs(f[0, v] - er(r) * xc / r) <= tol, (ix, iy)
        r = np.hypot(xl, yc)
        assert abs(f[1, v] - er(r) * yc / r) <= tol, (ix, iy)
    # every voxel (ghosts and all z layers) holds the same x-y field; B = (0, 0, bz)
    assert np.array_equal(f[0, :X * Y], f[0, 3 * X * Y:4 * X * Y])
    assert np.all(f[5] == bz) and not np.any(f[[2, 3, 4, 6, 7, 8]])
    p = d.load_particles()
    ix = p["cell"] % X
    iy = (p["cell"] // X) % Y
    xp = (ix - 1 + 0.5 * (p["x"] + 1)) * dx - L / 2
    yp = (iy - 1 + 0.5 * (p["y"] + 1)) * dx - L / 2
    r = np.hypot(xp, yp)
    th = np.arctan2(yp, xp)
    we = float(d.driver_constants()["we"])
    prof = we * np.exp(-((r - r0) / sig) ** 2) * (1 + pl["pert_amp"] * np.cos(3 * th))
    np.testing.assert_allclose(p["w"], prof, rtol=1e-12, atol=1e-300)
    sel = rng.choice(r.size, 60, replace=False)
    for i in sel:
        e = er(r[i])
        vx, vy = e * yp[i] / r[i] / bz, -e * xp[i] / r[i] / bz
        g = 1 / np.sqrt(1 - vx * vx - vy * vy)
        assert abs(p["ux"][i] - g * vx) <= tol / bz
        assert abs(p["uy"][i] - g * vy) <= tol / bz
    assert not np.any(p["uz"])
    # the drift is azimuthal, counter-clockwise about +z for an electron column
    assert np.all(np.abs(p["ux"] * xp + p["uy"] * yp) <= 1e-12 * (np.abs(p["ux"]) + np.abs(p["uy"])) * r)
    assert np.all(xp * p["uy"] - yp * p["ux"] >= -tol / bz * r)
    assert np.mean(xp * p["uy"] - yp * p["ux"]) > 0


# The reference's own deck sources (src/input/deck.h interface), compiled
# unchanged against cabanapic_amd/host through its ref_include/ path; read
# from the reference tree where it lies, so skipped where it is absent.
REF_ROOT = "/root/reference"
REF_DECKS = [("decks/2stream-short.cxx", "ref_2stream_short"),
             ("decks/2particle.cxx", "ref_2particle"),
             ("decks/custom_init.cxx", "ref_custom_init"),
             ("decks/dioctron_3d.cxx", "ref_dioctron_3d"),
             ("tests/energy_comparison/2stream-em.cxx", "ref_2stream_em")]


def _build_ref(src, name, real):
    subprocess.run(["make", "-s", "-C", host_driver.HOST, "ref_deck", f"SRC={REF_ROOT}/{src}",
                    f"NAME={name}", f"TYPE={real}"], check=True, capture_output=True)


@pytest.mark.skipif(not os.path.isdir(REF_ROOT), reason="reference tree not present")
@pytest.mark.parametrize("src,name", REF_DECKS, ids=[n for _, n in REF_DECKS])
def test_reference_deck_source_builds_unchanged(src, name, tmp_path):
    """Each deck file of the reference compiles unchanged against the deck
    interface (its Cabana::slice / KOKKOS_LAMBDA / simd_parallel_for
    initializers included) and initializes particles in interior cells."""
    _build_ref(src, name, "double")
    k, p = host_driver.dump(name, "double", str(tmp_path / "init.bin"))
    n = p["cell"].size
    assert n > 0 and k["num_steps"] > 0
    for m in ("x", "y", "z"):
        assert np.all(np.abs(p[m]) <= 1), m


@pytest.mark.skipif(not os.path.isdir(REF_ROOT), reason="reference tree not present")
@pytest.mark.parametrize("real", ["float", "double"])
def test_reference_regression_deck_equals_ours(real, tmp_path):
    """The reference's tests/energy_comparison/2stream-em.cxx, built
    unchanged, gives the same driver constants and initial particles, bit
    for bit, as this repository's regression deck."""
    _build_ref("tests/energy_comparison/2stream-em.cxx", "ref_2stream_em", real)
    k1, p1 = host_driver.dump("ref_2stream_em", real, str(tmp_path / "a.bin"))
    k2, p2 = host_driver.dump("two_stream_em_regression", real, str(tmp_path / "b.bin"))
    assert k1 == k2
    ui = np.uint32 if real == "float" else np.uint64
    for m in ("x", "y", "z", "ux", "uy", "uz", "w"):
        assert np.array_equal(p1[m].view(ui), p2[m].view(ui)), m
    assert np.array_equal(p1["cell"], p2["cell"])
