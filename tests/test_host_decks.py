"""CPU checks of the C++ deck layer (cabanapic_amd/host): the regression
deck's derived constants and default particle initializer reproduce the
oracle's restatement of src/input/deck.h and example/example.cpp bit for bit,
and the Python decks agree with the C++ decks on the derived constants."""
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
                                          ("two_stream_3d", decks.two_stream_3d)])
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
