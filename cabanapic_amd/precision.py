"""FP32 against FP64 on the Weibel deck.

The same initial plasma (the counter-based loader draws identical particles
in both precisions, up to the rounding of the FP32 copy) is run for a number
of steps in FP32, in FP32 with compensated current sums, and in FP64.  The
figure compared is the B-field energy history, EM_Field_Solver::b_energy
(src/fields.h:588-615), which for the Weibel instability grows exponentially
from noise and saturates (on 128^3 cells x 32 particles per cell, dx = 0.1
c/w_pe: noise until about step 300, growth until about step 800, the largest
W_B near step 2100), so the default run is 2500 steps; reported per FP32 run
against the FP64 run:

* the error of the growth curve: max over steps of |ln(W32 / W64)| over the
  whole run, over its first `first` steps and over the linear-growth window,
  and the max relative error over the window;
* the growth rate: slope of 0.5 ln W_B against time (the field amplitude's
  e-folding rate) by least squares over the linear-growth window;
* the saturation: the largest W_B of the run and the step it is reached at.

The linear-growth window is taken from the FP64 curve: the steps before its
maximum whose ln W_B lies in the middle half of the range between the noise
level (mean of ln W_B over the first 10 steps) and the maximum.
"""
import time

import numpy as np

from . import decks
from .pic import Simulation


def growth_window(b):
    """(lo, hi) step indices of the linear-growth window of the curve b."""
    lb = np.log(np.maximum(b, np.finfo(np.float64).tiny))
    top = int(np.argmax(b))
    noise = float(lb[:10].mean())
    span = lb[top] - noise
    lo_v, hi_v = noise + 0.25 * span, noise + 0.75 * span
    idx = np.nonzero((lb[:top + 1] >= lo_v) & (lb[:top + 1] <= hi_v))[0]
    if idx.size < 2:
        raise ValueError("no linear-growth window: the curve does not grow")
    return int(idx[0]), int(idx[-1]) + 1


def growth_rate(b, dt, window):
    lo, hi = window
    t = dt * np.arange(lo, hi)
    return float(np.polyfit(t, 0.5 * np.log(b[lo:hi]), 1)[0])


def run(deck, steps, compensated=False):
    """E and B energy history (steps x 2) of the deck, and the seconds the
    steps took on the device."""
    parts = deck.load_particles()
    sim = Simulation(deck.nx, deck.ny, deck.nz, deck.num_particles, deck.R)
    try:
        sim.upload_particles(*[parts[k] for k in ("x", "y", "z", "ux", "uy", "uz", "w")],
                             parts["cell"])
        del parts
        sim.set_compensated_deposit(compensated)
        sp = deck.step_params(energy_interval=1)
        sim.synchronize()
        t0 = time.perf_counter()
        h = sim.step(sp, steps)
        sim.synchronize()
        t1 = time.perf_counter()
        if sim.device_errors():
            raise RuntimeError(f"device error flags {sim.device_errors()}")
    finally:
        sim.close()
    return np.asarray(h)[:, :2].copy(), t1 - t0


def compare(b32, b64, dt, first=500):
    """The figures of one FP32 B-energy curve against the FP64 one."""
    win = growth_window(b64)
    lo, hi = win
    pos = (b32 > 0) & (b64 > 0)
    lr = np.abs(np.log(b32[pos] / b64[pos]))
    f = min(first, b64.size)
    lr_first = np.abs(np.log(b32[:f] / b64[:f]))
    g32, g64 = growth_rate(b32, dt, win), growth_rate(b64, dt, win)
    i32, i64 = int(np.argmax(b32)), int(np.argmax(b64))
    return {
        "window_steps": [lo + 1, hi],
        "curve_max_abs_log_err": float(lr.max()),
        "curve_first_steps": f,
        "curve_max_abs_log_err_first": float(lr_first.max()),
        "curve_max_abs_log_err_window": float(np.abs(np.log(b32[lo:hi] / b64[lo:hi])).max()),
        "curve_max_rel_err_window": float((np.abs(b32[lo:hi] - b64[lo:hi]) / b64[lo:hi]).max()),
        "growth_rate": g32, "growth_rate_fp64": g64,
        "growth_rate_rel_err": abs(g32 - g64) / abs(g64),
        "saturation": float(b32[i32]), "saturation_fp64": float(b64[i64]),
        "saturation_rel_err": abs(float(b32[i32]) - float(b64[i64])) / float(b64[i64]),
        "saturation_step": i32 + 1, "saturation_step_fp64": i64 + 1,
    }


def weibel_report(steps=2500, n=128, nppc=32, first=500):
    """Runs the Weibel deck (n^3 cells, nppc particles per cell) for `steps`
    steps in FP64, FP32 and compensated FP32; returns the histories and the
    figures of both FP32 runs against FP64 (the curve error also over the
    first `first` steps), with their ms per step."""
    d64 = decks.weibel_3d(np.float64, n=n, nppc=nppc)
    d32 = decks.weibel_3d(np.float32, n=n, nppc=nppc)
    h64, t64 = run(d64, steps)
    h32, t32 = run(d32, steps)
    h32c, t32c = run(d32, steps, compensated=True)
    dt = float(d64.dt)
    plain = compare(h32[:, 1], h64[:, 1], dt, first)
    comp = compare(h32c[:, 1], h64[:, 1], dt, first)
    plain["ms_per_step"] = 1e3 * t32 / steps
    comp["ms_per_step"] = 1e3 * t32c / steps
    return {"steps": steps, "deck": d64.name, "dt": dt, "fp64_ms_per_step": 1e3 * t64 / steps,
            "fp32": plain, "fp32_compensated": comp,
            "history": {"fp64": h64, "fp32": h32, "fp32_compensated": h32c}}
