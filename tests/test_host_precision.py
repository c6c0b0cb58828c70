"""The growth-curve figures of cabanapic_amd.precision on closed-form curves:
noise, then W = A exp(2 g t), then a plateau.  The fitted rate is g, the
window lies inside the exponential phase, the saturation is the plateau's
first step, and a curve scaled by (1 + d) has |ln ratio| = ln(1 + d)
everywhere and the same rate."""
import math

import numpy as np

from cabanapic_amd import precision


def _curve(g=0.1, dt=0.05, n=3000, t0=300, t1=1500, noise=2.0):
    t = dt * np.arange(n)
    b = np.full(n, noise)
    grow = noise * np.exp(2 * g * (t - t[t0]))
    b[t0:t1] = grow[t0:t1]
    b[t1:] = grow[t1 - 1]
    return b, dt


def test_growth_rate_and_window_of_exponential():
    b, dt = _curve()
    lo, hi = precision.growth_window(b)
    assert 300 <= lo < hi <= 1500
    assert abs(precision.growth_rate(b, dt, (lo, hi)) - 0.1) < 1e-9


def test_compare_scaled_curve():
    b, dt = _curve()
    d = 1e-4
    f = precision.compare(b * (1 + d), b, dt, first=500)
    assert abs(f["curve_max_abs_log_err"] - math.log1p(d)) < 1e-12
    assert abs(f["curve_max_abs_log_err_first"] - math.log1p(d)) < 1e-12
    assert f["curve_first_steps"] == 500
    assert f["growth_rate_rel_err"] < 1e-9
    assert abs(f["saturation_rel_err"] - d) < 1e-12
    assert f["saturation_step"] == f["saturation_step_fp64"] == 1500


def test_compare_identical_and_shifted():
    b, dt = _curve()
    f = precision.compare(b.copy(), b, dt)
    assert f["curve_max_abs_log_err"] == 0 and f["growth_rate_rel_err"] == 0
    # a curve that grows 1% faster shows a 1% rate error
    b2, _ = _curve(g=0.101)
    f = precision.compare(b2, b, dt)
    assert abs(f["growth_rate_rel_err"] - 0.01) < 1e-6


def test_flat_curve_has_no_window():
    try:
        precision.growth_window(np.ones(100))
    except ValueError:
        return
    raise AssertionError("a flat curve has no growth window")
