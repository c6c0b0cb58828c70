#!/usr/bin/env python3
"""Throughput of the PIC step on one MI355X: particle-steps/s.

Workload: the 128^3 x 32 ppc Weibel-type configuration (67,108,864
electrons, bi-Maxwellian u_th = (0.3, 0.05, 0.05) c, seed 11, dx = 0.1,
dt = 0.99 Courant), FP32 push/deposit.  A step is one pass of the reference
loop body (example/example.cpp:216-279): interpolator load, accumulator
clear, fused push + move_p + deposit, unload, advance_b/2, advance_e,
advance_b/2, FP64 field-energy reduction.  Inputs are resident in HBM before
the timed region.  The FP64 path is timed on the same workload for the
FP32/FP64 ratio, and the CPU baseline is the oracle (OpenMP) on a bounded
number of steps of the same workload.

Prints one JSON line.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD published


def algorithmic_bytes(np_, nc, itemsize):
    """Per-step algorithmic HBM bytes: particles read x,y,z,ux,uy,uz,w,cell
    and write x,y,z,ux,uy,uz; per interior cell E,B read+write, J write+read,
    energy read of E,B."""
    per_p = 7 * itemsize + 4 + 6 * itemsize
    per_c = 24 * itemsize
    return per_p * np_, per_c * nc


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "push_stage_weibel128x32_traffic.md")


def measured_push_traffic(n, nppc):
    """HBM bytes per push stage from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (scaled by the copy calibration), for the workload it
    was measured on; None for any other workload."""
    if (n, nppc) != (128, 32) or not os.path.exists(TRAFFIC_PROFILE):
        return None
    with open(TRAFFIC_PROFILE) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_push_stage:"):
                return float(line.split(":")[1])
    return None


COUNTERS_PROFILE = os.path.join(ROOT, "profiles", "push_stage_weibel128x32_counters.md")


def measured_valu_busy(n, nppc):
    """{kernel: share of its cycles issuing VALU work} from the committed
    rocprofv3 SQ counter summary of the push passes, for the workload it was
    measured on; None for any other workload."""
    if (n, nppc) != (128, 32) or not os.path.exists(COUNTERS_PROFILE):
        return None
    out = {}
    with open(COUNTERS_PROFILE) as f:
        for line in f:
            if line.startswith("| `cpic::k_") and line.rstrip().endswith("% |"):
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                name = cells[0].strip("`").split("::")[1].split("<")[0]
                out[name] = round(float(cells[-1].rstrip(" %")) / 100, 3)
    return out or None


def copy_bandwidth(reps=10):
    """Achieved bandwidth of a plain 1 GiB device-to-device copy (torch),
    GB/s counting bytes read plus bytes written: the practical streaming
    ceiling of the box beside the published 8 TB/s.  Runs in a process of
    its own (`--copy-bandwidth`): torch's HIP runtime does not start in a
    process where the library's has."""
    import torch
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    torch.cuda.empty_cache()
    return 2 * 4 * n / (ms * 1e-3) / 1e9


def measured_copy_bandwidth(reps):
    """copy_bandwidth(reps) from a child process; None if it fails."""
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--copy-bandwidth",
                        "--copy-reps", str(reps)],
                       capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        print(f"# copy bandwidth run failed: {r.stderr.strip()[-300:]}", file=sys.stderr)
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])["copy_gbs"]


def run_gpu(deck, steps, warmup, compact_interval, history):
    from cabanapic_amd import Simulation
    parts = deck.load_particles()
    sim = Simulation(deck.nx, deck.ny, deck.nz, deck.num_particles, deck.R)
    tu = time.perf_counter()
    sim.upload_particles(*[parts[k] for k in ("x", "y", "z", "ux", "uy", "uz", "w")],
                         parts["cell"])
    t_upload = time.perf_counter() - tu
    del parts
    fields = deck.load_fields()
    if fields is not None:
        sim.upload_fields(fields)
        del fields
    sp = deck.step_params(energy_interval=1, compact_interval=compact_interval)
    # Gauss's law baseline (FP64 on the device), before any step
    gargs = (float(deck.dx), float(deck.dy), float(deck.dz), float(deck.eps), float(deck.qsp))
    sim.gauss_residual(*gargs, baseline=True)
    if warmup:
        sim.step(sp, warmup)
    # inside the timed region only the push stage is bracketed by events
    sim.set_timing(True, stages=("push",))
    sim.synchronize()
    t0 = time.perf_counter()
    hist = sim.step(sp, steps, history=history)
    sim.synchronize()
    t1 = time.perf_counter()
    stages = sim.stage_times()
    # the other stages' times from a few more steps, outside the timed region
    sim.set_timing(True)
    sim.step(sp, 5, history=False)
    for name, v in sim.stage_times().items():
        if name != "push":
            stages[name] = v
    change, rho = sim.gauss_residual(*gargs, baseline=False)
    stages["gauss_residual_change_rel"] = change / rho
    errs = sim.device_errors()
    td = time.perf_counter()
    sim.download_particles()
    stages["host_transfer_s"] = (t_upload, time.perf_counter() - td)
    sim.close()
    return t1 - t0, hist, stages, errs


def run_tangent(deck, steps, warmup):
    """The tangent-linear step (derivative with respect to a drift of ux)
    on the same workload; returns seconds for `steps` steps."""
    from cabanapic_amd import TangentSimulation
    parts = deck.load_particles()
    n = deck.num_particles
    tan = np.zeros((6, n))
    tan[3] = 1.0
    ts = TangentSimulation(deck.nx, deck.ny, deck.nz, n, deck.R)
    ts.upload_particles(*[parts[k] for k in ("x", "y", "z", "ux", "uy", "uz", "w")],
                        parts["cell"], tan)
    del parts, tan
    sp = deck.step_params(energy_interval=1)
    if warmup:
        ts.step(sp, warmup)
    ts.primal.synchronize()
    t0 = time.perf_counter()
    hist = ts.step(sp, steps)
    t1 = time.perf_counter()
    errs = ts.primal.device_errors()
    ts.close()
    if errs or not np.isfinite(hist).all():
        raise SystemExit(f"tangent step failed (device error flags {errs})")
    return t1 - t0


def differentiation_config_costs(steps):
    """Forward step and tangent-linear step on the differentiation
    configuration (3D two-stream, 32^3 cells, 2 x 16 ppc, FP32 primal, FP64
    tangent): ms per step of each."""
    from cabanapic_amd import decks
    d = decks.two_stream_3d(np.float32, n=32, nppc=32, steps=steps)
    t_fwd, _, _, errs = run_gpu(d, steps, 10, 0, True)
    if errs:
        raise SystemExit(f"device error flags {errs}")
    t_tan = run_tangent(d, steps, 10)
    return 1e3 * t_fwd / steps, 1e3 * t_tan / steps


def adjoint_costs(steps, interval):
    """Reverse-mode derivative on the differentiation configuration (3D
    two-stream, 32^3 cells, 2 x 16 ppc, FP32 primal, FP64 adjoint) of the
    final E energy: ms per step of the plain forward run (cpic_step), of
    cpic_forward with a checkpoint every `interval` steps, and of
    cpic_backward (recompute and transpose), all from the same initial
    state, and the gradient with respect to the beam drift."""
    from cabanapic_amd import Adjoint, Simulation, decks
    d = decks.two_stream_3d(np.float32, n=32, nppc=32, steps=steps)
    parts = d.load_particles()
    sp = d.step_params(energy_interval=1)
    members = ("x", "y", "z", "ux", "uy", "uz", "w")
    sim = Simulation(d.nx, d.ny, d.nz, d.num_particles, np.float32)

    def upload():
        sim.upload_particles(*[parts[m] for m in members], parts["cell"])
        sim.upload_fields(np.zeros((9, d.num_cells), np.float32))

    upload()
    sim.step(sp, 10)  # warm-up
    upload()
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step(sp, steps)
    t_plain = time.perf_counter() - t0
    upload()
    with Adjoint(sim) as adj:
        t0 = time.perf_counter()
        adj.forward(sp, steps, interval)
        t_fwd = time.perf_counter() - t0
        f = sim.download_fields().astype(np.float64)
        seed = np.zeros_like(f)
        inner = np.zeros((d.nz + 2, d.ny + 2, d.nx + 2), bool)
        inner[1:-1, 1:-1, 1:-1] = True
        seed[:3, inner.ravel()] = f[:3, inner.ravel()]
        t0 = time.perf_counter()
        _, gp = adj.backward(seed)
        t_bwd = time.perf_counter() - t0
        mem = adj.memory()
    errs = sim.device_errors()
    sim.close()
    if errs:
        raise SystemExit(f"adjoint run: device error flags {errs}")
    # d(ux)/d(v0) of the loader: ux = s d(v0) (1 + a sin) + uth n, linear in the drift
    rel = 1e-6
    lo = decks.two_stream_3d(np.float64, n=32, nppc=32, steps=steps, v0=0.2 * (1 - rel))
    hi = decks.two_stream_3d(np.float64, n=32, nppc=32, steps=steps, v0=0.2 * (1 + rel))
    dux = (hi.load_particles()["ux"] - lo.load_particles()["ux"]) / (2 * rel * 0.2)
    grad = float(np.dot(gp[3], dux))
    return (1e3 * t_plain / steps, 1e3 * t_fwd / steps, 1e3 * t_bwd / steps, grad, mem)


def column_gradient(steps, interval):
    """dA_3/dB_z on the 192 x 192 x 64 x 32 electron column: the m = 3
    amplitude of the radial E field after `steps` FP32 steps, differentiated
    with respect to the guide field by cpic_forward / cpic_backward with a
    checkpoint every `interval` steps."""
    from cabanapic_amd import Adjoint, Simulation, decks
    d = decks.dioctron_3d(np.float32, steps=steps)
    parts = d.load_particles()
    sim = Simulation(d.nx, d.ny, d.nz, d.num_particles, np.float32)
    sim.upload_particles(*[parts[m] for m in ("x", "y", "z", "ux", "uy", "uz", "w")],
                         parts["cell"])
    del parts
    sim.upload_fields(d.load_fields())
    sp = d.step_params(energy_interval=1)
    with Adjoint(sim) as adj:
        t0 = time.perf_counter()
        adj.forward(sp, steps, interval)
        t_fwd = time.perf_counter() - t0
        amp, seed = decks.column_mode_amplitude(d, sim.download_fields())
        t0 = time.perf_counter()
        gf, gp = adj.backward(seed)
        t_bwd = time.perf_counter() - t0
        mem = adj.memory()
    errs = sim.device_errors()
    sim.close()
    if errs:
        raise SystemExit(f"column adjoint: device error flags {errs}")
    tf, tp = decks.column_guide_field_tangent(d)
    grad = float(np.vdot(gf, tf) + np.vdot(gp, tp))
    return amp, grad, 1e3 * t_fwd / steps, 1e3 * t_bwd / steps, mem


def cpu_description():
    """CPU model and physical core count of this host (from /proc/cpuinfo),
    and the cores this process may run on."""
    model, cores = "unknown", set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model == "unknown":
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None and core is not None:
                    cores.add((phys, core))
                    phys = core = None
        if phys is not None and core is not None:
            cores.add((phys, core))
    except OSError:
        pass
    allowed = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return model, len(cores) or (os.cpu_count() or 1), allowed


def cpu_threads():
    """Threads of the CPU baseline: one per physical core the process may
    use, at most OMP_NUM_THREADS where the environment sets it (the CPU share
    of one GPU on a shared host)."""
    _, phys, allowed = cpu_description()
    n = min(phys, allowed)
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(n, 1)


def cpu_baseline(deck, nsteps, nthreads):
    """The oracle's OpenMP step on the full workload for nsteps steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    parts = deck.load_particles()
    st = oracle_lib.OracleState(deck.nx, deck.ny, deck.nz, deck.num_particles, deck.R)
    for k in ("x", "y", "z", "ux", "uy", "uz", "w"):
        st.p[k] = parts[k]
    st.cell = parts["cell"]
    del parts
    k = deck.step_params()
    consts = [k[n] for n in ("qdt_2mc", "cdt_dx", "cdt_dy", "cdt_dz", "qsp", "dx", "dy", "dz",
                             "dt", "px", "py", "pz", "dt_eps0", "me", "c")]
    st.run(consts, 1, nthreads=nthreads)  # first touch of the arrays
    t0 = time.perf_counter()
    st.run(consts, nsteps, nthreads=nthreads)
    dt = time.perf_counter() - t0
    return deck.num_particles * nsteps / dt, dt


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64,
                    help="timed steps (64: two full compaction intervals of 32 pushes)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=128, help="cells per axis")
    ap.add_argument("--nppc", type=int, default=32)
    ap.add_argument("--compact-interval", type=int, default=0,
                    help="steps between column compactions (0 = library default)")
    ap.add_argument("--fp64-steps", type=int, default=20,
                    help="steps of the FP64 path for the FP32/FP64 ratio (0 = skip)")
    ap.add_argument("--cpu-steps", type=int, default=30,
                    help="steps of the CPU baseline (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU baseline (0 = one per physical core the process "
                         "may use, within OMP_NUM_THREADS)")
    ap.add_argument("--tangent-steps", type=int, default=5,
                    help="timed steps of the tangent-linear step (0 = skip)")
    ap.add_argument("--adjoint-steps", type=int, default=500,
                    help="steps of the reverse-mode run on the differentiation configuration "
                         "(0 = skip)")
    ap.add_argument("--column-gradient-steps", type=int, default=1000,
                    help="steps of the electron column's dA_3/dB_z adjoint run (0 = skip)")
    ap.add_argument("--precision-steps", type=int, default=2500,
                    help="steps of the FP32 / compensated FP32 / FP64 Weibel runs whose "
                         "B-energy growth curves are compared, to past the saturation "
                         "(0 = skip)")
    ap.add_argument("--column-steps", type=int, default=20,
                    help="timed FP32 steps of the 192x192x64 x 32 ppc electron-column "
                         "configuration, reported beside the metric (0 = skip)")
    ap.add_argument("--copy-reps", type=int, default=10,
                    help="timed repetitions of the 1 GiB device copy reported as copy_gbs "
                         "(0 = skip, e.g. under a profiler)")
    ap.add_argument("--copy-bandwidth", action="store_true",
                    help="only time a plain 1 GiB device-to-device copy (torch) and print "
                         "its GB/s (read plus written bytes)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (torch), the known-size "
                         "kernel that scales rocprofv3 FETCH_SIZE/WRITE_SIZE to bytes")
    args = ap.parse_args()
    # thread placement of the CPU baseline: set before the OpenMP runtime
    # starts (the oracle library loads it)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")

    if args.copy_bandwidth:
        print(json.dumps({"copy_gbs": round(copy_bandwidth(args.copy_reps), 1)}))
        return

    if args.copy_calibration:
        import torch
        n = 1 << 28  # 2^28 float32 = 1 GiB
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}))
        return

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): the step has no "
              "multi-GPU path; its domain decomposition is a design, not code.", file=sys.stderr)
        sys.exit(2)

    # before the library starts its HIP runtime in this process
    copy_gbs = measured_copy_bandwidth(args.copy_reps) if args.copy_reps > 0 else None

    from cabanapic_amd import decks

    d32 = decks.weibel_3d(np.float32, n=args.n, nppc=args.nppc)
    np_, nc = d32.num_particles, d32.nx * d32.ny * d32.nz
    print(f"# workload {d32.name}: {np_} particles, {nc} cells, dt={float(d32.dt):.6g}",
          file=sys.stderr, flush=True)

    t32, h32, st32, err32 = run_gpu(d32, args.steps, args.warmup, args.compact_interval, True)
    if err32:
        raise SystemExit(f"device error flags {err32}")
    value = np_ * args.steps / t32
    ms_step = 1e3 * t32 / args.steps
    print(f"# fp32: {ms_step:.3f} ms/step, {value:.4g} particle-steps/s", file=sys.stderr,
          flush=True)

    bp, bc = algorithmic_bytes(np_, nc, 4)
    push_n, push_ms = st32["push"]
    push_avg_s = push_ms / max(push_n, 1) * 1e-3
    push_gbs = bp / push_avg_s / 1e9
    step_gbs = (bp + bc) / (t32 / args.steps) / 1e9

    gauss32 = st32.pop("gauss_residual_change_rel")
    t_up, t_down = st32.pop("host_transfer_s")
    extra = {"stage_ms_per_step": {k: round(v[1] / max(v[0], 1), 4) for k, v in st32.items()
                                   if v[0]},
             "gauss_residual_change_rel": gauss32,
             "step_hbm_gbs": round(step_gbs, 1),
             "copy_gbs": copy_gbs,
             # the push passes are co-bound by VALU issue (SQ counters)
             "push_valu_busy": measured_valu_busy(args.n, args.nppc),
             # particles handed over in host buffers: upload (binning into
             # columns included) before and download after the timed steps
             "upload_s": round(t_up, 4), "download_s": round(t_down, 4),
             "pcie_inclusive_value": np_ * args.steps / (t_up + t32 + t_down),
             "step_roofline_frac": round(step_gbs / HBM_PEAK_GBS, 4)}

    if args.fp64_steps > 0:
        d64 = decks.weibel_3d(np.float64, n=args.n, nppc=args.nppc)
        nwin = min(args.fp64_steps, args.steps)
        t64, h64, st64, err64 = run_gpu(d64, args.fp64_steps, args.warmup, args.compact_interval,
                                        True)
        v64 = np_ * args.fp64_steps / t64
        gauss64 = st64.pop("gauss_residual_change_rel")
        print(f"# fp64: {1e3 * t64 / args.fp64_steps:.3f} ms/step, {v64:.4g} particle-steps/s",
              file=sys.stderr, flush=True)
        # field energy of the FP32 run against the FP64 run of the same deck,
        # over the steps both timed (same warmup, so the same step numbers)
        e32 = h32[:nwin, 0] + h32[:nwin, 1]
        e64 = h64[:nwin, 0] + h64[:nwin, 1]
        rel = np.abs(e32 - e64) / np.abs(e64)
        extra.update(fp64_value=v64, fp32_over_fp64=round(value / v64, 3),
                     fp64_gauss_residual_change_rel=gauss64,
                     field_energy_rel_err_vs_fp64=float(rel.max()),
                     field_energy_window=f"steps {args.warmup + 1}-{args.warmup + nwin}")

    if args.tangent_steps > 0:
        tt = run_tangent(d32, args.tangent_steps, 1)
        tms = 1e3 * tt / args.tangent_steps
        print(f"# tangent fp32: {tms:.3f} ms/step ({tms / ms_step:.2f} x the step)",
              file=sys.stderr, flush=True)
        extra.update(tangent_ms_per_step=round(tms, 3),
                     tangent_over_forward=round(tms / ms_step, 3))
        f4, t4 = differentiation_config_costs(100)
        print(f"# differentiation config 32^3 x 32: forward {f4:.3f} ms/step, tangent "
              f"{t4:.3f} ms/step ({t4 / f4:.2f} x)", file=sys.stderr, flush=True)
        extra.update(diffconfig_forward_ms_per_step=round(f4, 4),
                     diffconfig_tangent_ms_per_step=round(t4, 4),
                     diffconfig_tangent_over_forward=round(t4 / f4, 3))

    if args.adjoint_steps > 0:
        k = 50
        f_ms, af_ms, ab_ms, g, mem = adjoint_costs(args.adjoint_steps, k)
        print(f"# adjoint, differentiation config 32^3 x 32, {args.adjoint_steps} steps, "
              f"checkpoints every {k}: forward {f_ms:.3f} ms/step, forward with checkpoints "
              f"{af_ms:.3f}, backward {ab_ms:.3f} ({(af_ms + ab_ms) / f_ms:.2f} x the forward "
              f"run); dW_E/dv0 = {g:.6e}", file=sys.stderr, flush=True)
        extra.update(adjoint_diffconfig_steps=args.adjoint_steps,
                     adjoint_diffconfig_checkpoint_interval=k,
                     adjoint_diffconfig_forward_ms_per_step=round(f_ms, 4),
                     adjoint_diffconfig_checkpointed_forward_ms_per_step=round(af_ms, 4),
                     adjoint_diffconfig_backward_ms_per_step=round(ab_ms, 4),
                     adjoint_diffconfig_backward_over_forward=round(ab_ms / f_ms, 3),
                     adjoint_diffconfig_total_over_forward=round((af_ms + ab_ms) / f_ms, 3),
                     adjoint_diffconfig_dWE_dv0=g,
                     adjoint_diffconfig_hbm_bytes=mem)

    if args.column_gradient_steps > 0:
        k = max(1, int(round(args.column_gradient_steps ** 0.5)))
        amp, g, cf_ms, cb_ms, mem = column_gradient(args.column_gradient_steps, k)
        print(f"# electron column dA_3/dB_z after {args.column_gradient_steps} FP32 steps "
              f"(checkpoints every {k}): A_3 = {amp:.6e}, dA_3/dB_z = {g:.6e}; forward "
              f"{cf_ms:.3f} ms/step, backward {cb_ms:.3f} ms/step; HBM {mem}", file=sys.stderr,
              flush=True)
        extra.update(column_gradient_steps=args.column_gradient_steps,
                     column_gradient_checkpoint_interval=k,
                     column_mode3_amplitude=amp, column_dA3_dBz=g,
                     column_gradient_forward_ms_per_step=round(cf_ms, 4),
                     column_gradient_backward_ms_per_step=round(cb_ms, 4),
                     column_gradient_hbm_bytes=mem)

    if args.precision_steps > 0:
        from cabanapic_amd import precision
        r = precision.weibel_report(args.precision_steps, n=args.n, nppc=args.nppc)
        keys = ("window_steps", "curve_max_abs_log_err", "curve_first_steps",
                "curve_max_abs_log_err_first", "curve_max_abs_log_err_window",
                "curve_max_rel_err_window", "growth_rate", "growth_rate_rel_err", "saturation",
                "saturation_rel_err", "saturation_step", "ms_per_step")
        rep = {name: {k: r[name][k] for k in keys} for name in ("fp32", "fp32_compensated")}
        rep["fp64"] = {k: r["fp32"][k + "_fp64"] for k in ("growth_rate", "saturation",
                                                             "saturation_step")}
        rep["fp64"]["ms_per_step"] = r["fp64_ms_per_step"]
        for name in ("fp32", "fp32_compensated"):
            f = rep[name]
            print(f"# {args.precision_steps} Weibel steps, {name} vs fp64: growth rate "
                  f"{f['growth_rate']:.5f} ({f['growth_rate_rel_err']:.2e} rel.), saturation "
                  f"{f['saturation']:.4e} at step {f['saturation_step']} "
                  f"({f['saturation_rel_err']:.2e} rel.), curve |ln ratio| max "
                  f"{f['curve_max_abs_log_err']:.2e} ({f['curve_max_abs_log_err_first']:.2e} "
                  f"over the first {f['curve_first_steps']} steps, "
                  f"{f['curve_max_abs_log_err_window']:.2e} in the growth window); "
                  f"{f['ms_per_step']:.3f} ms/step", file=sys.stderr,
                  flush=True)
        extra.update(precision_weibel_steps=args.precision_steps, precision_weibel=rep,
                     compensated_over_plain_step=round(rep["fp32_compensated"]["ms_per_step"] /
                                                       rep["fp32"]["ms_per_step"], 3))

    if args.column_steps > 0:
        dc = decks.dioctron_3d(np.float32)
        tc, _, stc, errc = run_gpu(dc, args.column_steps, args.warmup, args.compact_interval,
                                   True)
        extra.update(column_gauss_residual_change_rel=stc.pop("gauss_residual_change_rel"))
        if errc:
            raise SystemExit(f"device error flags {errc}")
        vc = dc.num_particles * args.column_steps / tc
        print(f"# {dc.name}: {1e3 * tc / args.column_steps:.3f} ms/step, {vc:.4g} "
              "particle-steps/s", file=sys.stderr, flush=True)
        extra.update(column_workload=dc.name, column_particles=dc.num_particles,
                     column_value=vc, column_ms_per_step=round(1e3 * tc / args.column_steps, 4))

    cpu = None
    if args.cpu_steps > 0:
        nth = args.cpu_threads or cpu_threads()
        model, phys, allowed = cpu_description()
        v_cpu, t_cpu = cpu_baseline(d32, args.cpu_steps, nth)
        n64 = max(1, args.cpu_steps // 2)
        v_cpu64, t_cpu64 = cpu_baseline(decks.weibel_3d(np.float64, n=args.n, nppc=args.nppc),
                                         n64, nth)
        cpu = {"value": v_cpu, "unit": "particle-steps/s", "cores": nth, "kind": "port",
               "sample": f"oracle OpenMP step (C restatement of the reference step, not the "
                         f"Kokkos build; per-thread accumulator copies summed after the push) "
                         f"on the full {d32.name} workload, FP32, {args.cpu_steps} steps, "
                         f"{t_cpu:.1f} s",
               "fp64_value": v_cpu64,
               "fp64_sample": f"same, FP64, {n64} steps, {t_cpu64:.1f} s",
               "cpu_model": model, "physical_cores": phys, "cores_allowed": allowed,
               "binding": f"OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')} "
                          f"OMP_PLACES={os.environ.get('OMP_PLACES')}"}
        print(f"# cpu ({model}, {phys} physical cores, {allowed} allowed): FP32 {v_cpu:.4g}, "
              f"FP64 {v_cpu64:.4g} particle-steps/s on {nth} threads", file=sys.stderr,
              flush=True)

    line = {
        "metric": "particle-steps/s at 128^3x32ppc, FP32 vs plain FP64 (<=1e-3 rel energy err)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": d32.name, "cells": [d32.nx, d32.ny, d32.nz], "nppc": d32.nppc,
                   "particles": np_, "compact_interval": args.compact_interval,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(push_gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(push_gbs / HBM_PEAK_GBS, 4),
                     "kernel": "push stage: k_push_rows (k_push_pass every 32nd step) + "
                               "k_arrive_pass + 2 x k_finish_pass + k_spill_pass",
                     "traffic": measured_push_traffic(args.n, args.nppc),
                     "traffic_unit": "bytes per push stage (one step)",
                     "algorithmic_bytes": bp},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
