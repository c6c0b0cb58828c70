"""The reference's energy regression rule, restated for the tests.

tests/energy_comparison/compare_energies.h:55-106 (relative error against
min(A,B), absolute error below 2 eps when min(A,B) < 10 eps) and the deck's
use of it, tests/energy_comparison/2stream-em.cxx:45-71 (skip 3581 lines,
compare the next 1300, E column then B column, 10 % margin).
"""
import os

import numpy as np

GOLD_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EPS = np.finfo(np.float64).eps


def gold(precision):
    """[step, time, E, B] rows of the reference gold file ('float'/'double')."""
    return np.loadtxt(os.path.join(GOLD_DIR, "energies_gold.2stream-em." + precision))


def compare_error(a, b, tol):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    mn = np.minimum(a, b)
    small = np.abs(mn) < 10 * EPS
    err = np.where(small, np.abs(a - b), np.abs(a - b) / np.where(small, 1.0, mn))
    ok = np.where(small, err < 2 * EPS, err < tol)
    return ok, err


def reference_rule(e, b, precision, skip=3581, count=1300, tol=0.10):
    """Returns (passed, max_err_E, max_err_B) for energy columns e, b."""
    g = gold(precision)
    sl = slice(skip, skip + count)
    ok_e, err_e = compare_error(g[sl, 2], np.asarray(e)[sl], tol)
    ok_b, err_b = compare_error(g[sl, 3], np.asarray(b)[sl], tol)
    return bool(ok_e.all() and ok_b.all()), float(err_e.max()), float(err_b.max())


def growth_rate(b, dt, lo=1000, hi=3000):
    """Least-squares slope of ln W_B per unit time over rows [lo, hi)."""
    t = np.arange(lo, hi) * dt
    return float(np.polyfit(t, np.log(np.asarray(b)[lo:hi]), 1)[0])
