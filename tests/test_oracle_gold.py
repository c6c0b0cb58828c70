"""Pins the CPU oracle against the reference's own gold energy histories.

The 2stream-em deck (tests/energy_comparison/2stream-em.cxx:81-114) with the
default particle initializer (src/input/deck.h:79-158) is run for its full
6000 steps through the oracle's restatement of the step loop
(example/example.cpp:216-279), in both precisions, and compared with
energies_gold.2stream-em.{double,float} (copied to tests/golden/).
"""
import numpy as np
import pytest

import energy_rule

_cache = {}


def history(oracle, precision):
    if precision not in _cache:
        dtype = np.float64 if precision == "double" else np.float32
        d, st = oracle.setup_2stream_em(dtype)
        h, _ = st.run(d["consts"], d["num_steps"])
        _cache[precision] = (d, st, h)
    return _cache[precision]


@pytest.mark.parametrize("precision", ["double", "float"])
def test_reference_rule(oracle, precision):
    """The reference's pass rule: < 10 % on lines 3582-4881, E then B."""
    _, _, h = history(oracle, precision)
    # column 2/3: energies accumulated in real_t, as the reference reports them
    ok, ee, eb = energy_rule.reference_rule(h[:, 2], h[:, 3], precision)
    assert ok, (ee, eb)


def test_double_matches_gold_to_print_precision(oracle):
    """The double build is deterministic enough that the restated step must
    reproduce every printed digit (%e, 7 significant figures) of all 6000
    lines; a dropped term or a wrong sign anywhere in the step fails this."""
    _, _, h = history(oracle, "double")
    g = energy_rule.gold("double")
    rel_e = np.abs(h[:, 2] - g[:, 2]) / g[:, 2]
    rel_b = np.abs(h[:, 3] - g[:, 3]) / g[:, 3]
    assert rel_e.max() < 1e-6, rel_e.max()
    assert rel_b.max() < 1e-6, rel_b.max()


def test_float_early_steps_match_gold(oracle):
    """Before round-off is amplified the float build also matches its gold."""
    _, _, h = history(oracle, "float")
    g = energy_rule.gold("float")
    rel = np.abs(h[:10, 2:4] - g[:10, 2:4]) / g[:10, 2:4]
    assert rel.max() < 5e-6, rel.max()
    rel = np.abs(h[:1000, 3] - g[:1000, 3]) / g[:1000, 3]
    assert rel.max() < 1e-3, rel.max()


@pytest.mark.parametrize("precision", ["double", "float"])
def test_time_column_and_dt(oracle, precision):
    """Gold column 2 is (line-1)*dt, an int times real_t product evaluated
    in real_t (example/example.cpp:270), printed with %f."""
    d, _, _ = history(oracle, precision)
    g = energy_rule.gold(precision)
    rt = np.float64 if precision == "double" else np.float32
    t = (np.arange(g.shape[0]).astype(rt) * rt(d["dt"])).astype(np.float64)
    assert np.abs(g[:, 1] - t).max() < 6e-7
    assert abs(d["dt"] - 0.019549) < 5e-7


@pytest.mark.parametrize("precision", ["double", "float"])
def test_growth_rate(oracle, precision):
    """W_B grows as exp(2 gamma t) with 2 gamma = 0.2418 in both golds."""
    d, _, h = history(oracle, precision)
    rate = energy_rule.growth_rate(h[:, 3], d["dt"])
    gold_rate = energy_rule.growth_rate(energy_rule.gold(precision)[:, 3], d["dt"])
    assert abs(gold_rate - 0.2418) < 0.002
    assert abs(rate - gold_rate) / gold_rate < 0.01, (rate, gold_rate)


def test_fp64_and_real_energy_sums_agree(oracle):
    """The FP64 accumulation of the energies agrees with the real_t one."""
    _, _, h = history(oracle, "float")
    rel = np.abs(h[:, 0] - h[:, 2]) / h[:, 0]
    assert rel.max() < 1e-5


def test_particle_count_and_cells_stay_interior(oracle):
    d, st, _ = history(oracle, "double")
    nx, ny, nz = st.nx, st.ny, st.nz
    c = st.cell
    ix = c % (nx + 2)
    iy = (c // (nx + 2)) % (ny + 2)
    iz = c // ((nx + 2) * (ny + 2))
    assert c.size == d["np"]
    assert ix.min() >= 1 and ix.max() <= nx
    assert iy.min() >= 1 and iy.max() <= ny
    assert iz.min() >= 1 and iz.max() <= nz
    for k in ("x", "y", "z"):
        assert np.abs(st.p[k]).max() <= 1.0


def test_float_e_rule_is_order_sensitive(oracle):
    """In float the E window of the reference rule lies in the chaotic
    nonlinear phase: permuting the particles, which changes nothing but the
    summation order of the deposit, moves the E error across the 10 % line
    (seed 3: about 14 %, seed 0: about 5 %) while B stays below 1 %.  A float
    implementation with another deposit order is therefore held to the B rule
    and to the order-only E envelope, not to the E line itself."""
    errs = {}
    for seed in (0, 3):
        d, st = oracle.setup_2stream_em(np.float32)
        perm = np.random.default_rng(seed).permutation(st.np)
        for k in st.p:
            st.p[k][:] = st.p[k][perm]
        st.cell[:] = st.cell[perm]
        h, _ = st.run(d["consts"], d["num_steps"])
        ok, ee, eb = energy_rule.reference_rule(h[:, 2], h[:, 3], "float")
        errs[seed] = (ok, ee, eb)
        assert eb < 0.01, (seed, eb)
        assert ee < 0.20, (seed, ee)
    assert errs[0][0] and not errs[3][0], errs
