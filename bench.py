"""Throughput of the PowerLLEL channel-flow RK2 step on one MI355X.

Workload: the Re_tau~550 channel, 768 x 384 x 384 (2pi x pi x 2, tanh stretch
1.8, Re_b = 20000 -> re = 10000, dt = 5e-4), seeded log profile + 10% noise,
synthetic data.  One step = the full sequence of src/main.cc:116-154 (RK1,
halo/BC, RK2, halo/BC, Poisson RHS, FFT/tridiagonal Poisson solve, pressure
BC, projection, constant-flow forcing, halo/BC).

Metric: Mpoint-steps/s = nx*ny*nz*steps / seconds / 1e6, inputs resident in
HBM; three windows of --steps steps are timed and the median is reported.  The roofline entry reports the dominant kernel's algorithmic bytes per
launch over its average HIP-event duration against the 8 TB/s HBM peak.  The
CPU baseline is the FP64 oracle (the CPU restatement of the reference
algorithm) timed for a few steps (--cpu-steps, default six) of the same grid on
this host's cores (CPU model, thread count and visible cores are reported).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# name: (nx, ny, nz, lx, ly, lz, stretch, re, dt, profile)
WORKLOADS = {
    "retau550_768x384x384": (768, 384, 384, 6.283185307179586, 3.141592653589793, 2.0, 1.8,
                             10000.0, 5e-4, "log"),
    "retau180_256x192x128": (256, 192, 128, 12.566370614359172, 6.283185307179586, 2.0, 1.5,
                             2800.0, 1e-3, "poi"),
    "retau180_64x64x64": (64, 64, 64, 6.283185307179586, 3.141592653589793, 2.0, 1.5, 2800.0,
                          2e-3, "poi"),
    # the single-GPU memory-sizing case: 604 M points, ~60 GB of FP64 fields
    "retau1000_1536x768x512": (1536, 768, 512, 6.283185307179586, 3.141592653589793, 2.0, 2.0,
                               20000.0, 2.5e-4, "log"),
}
DEFAULT_WORKLOAD = "retau550_768x384x384"

# Algorithmic (compulsory) HBM bytes per grid point per launch.  Inside a run
# the previous step's projection is applied on load, so stage 1 reads u,v,w,p
# and writes u1,v1,w1.  FP64: stage 2 fused with calcRHS reads u1,v1,w1,u,v,w,p
# and writes u,v,w and the right-hand side; each transform and the tridiagonal
# solve read and write the work array once; correctVel (last step of a run)
# reads p,u,v,w and writes u,v,w (the bulk mean of u rides along); the separate
# bulk mean reads u; the force update reads and writes u.  FP32-mixed (4-byte
# fields): inside a run both stages run in one kernel (rk12: reads u,v,w,p,
# writes u,v,w; u1 stays on chip); the first and last step of a run take the
# separate stages, where stage 2 reads u1,v1,w1,u,v,w,p and writes u,v,w.
# calcRHS runs inside the forward x transform (reads u,v,w, writes the work
# array).
BYTES_PER_POINT = {
    "fp64": {"rk1": 56, "rk2": 72, "calc_rhs": 32, "rk2_rhs": 88, "fft_x_fwd": 16,
             "fft_y_fwd": 16, "solve_trid": 16, "fft_y_bwd": 16, "fft_x_bwd": 16,
             "correct_vel": 56, "mean_partial": 8, "force_add": 16},
    "fp32_mixed": {"rk1": 28, "rk2": 40, "rk12": 28, "calc_rhs": 16, "fft_x_fwd": 8,
                   "fft_x_fwd_rhs": 16,
                   "fft_y_fwd": 8, "solve_trid": 8, "fft_y_bwd": 8, "fft_x_bwd": 8,
                   "correct_vel": 28, "mean_partial": 4, "force_add": 8},
}
# Timed windows of --steps steps each; the reported time is their median.
WINDOWS = 3

# The compulsory bytes per point of each routine of the reference's RK2 step
# (src/main.cc:116-154, FP64, one launch each): their sum is the roofline of the
# whole step; FP32-mixed halves it.
ROUTINE_BYTES_FP64 = {"rk1": 48, "rk2": 72, "calc_rhs": 32, "fft_x_fwd": 16, "fft_y_fwd": 16,
                      "solve_trid": 16, "fft_y_bwd": 16, "fft_x_bwd": 16, "correct_vel": 56,
                      "mean_partial": 8, "force_add": 16}
STEP_BYTES_FP64 = 312

# HBM bytes per point per launch measured with rocprofv3 on the 768x384x384
# workload (2 x FETCH_SIZE + WRITE_SIZE, the scale of a 1 GiB device copy);
# profiles/retau550_768x384x384_step_kernels.md holds the runs.  Reported as
# the roofline entry's `traffic` for that workload only.
MEASURED_HBM_B_PER_POINT = {
    "fp64": {"rk1": 61.4, "rk2_rhs": 93.9, "fft_x_fwd": 16.0, "fft_y_fwd": 16.0, "solve_trid": 16.2,
             "fft_y_bwd": 16.0, "fft_x_bwd": 16.0, "correct_vel": 59.8},
    "fp32_mixed": {"rk1": 27.2, "rk2": 39.0, "rk12": 32.2, "fft_x_fwd_rhs": 20.7, "fft_y_fwd": 8.0,
                   "solve_trid": 8.2, "fft_y_bwd": 8.0, "fft_x_bwd": 8.0, "correct_vel": 31.6},
}

# MFMA flops per training sample of the closure network (forward 2(16*64 +
# 64*64), backward dW2, delta1 and dW1: 2(64*64 + 64*64 + 16*64)).
CLOSURE_TRAIN_FLOP = 10240 + 18432


def parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD, choices=sorted(WORKLOADS))
    ap.add_argument("--fp64-only", action="store_true", help="skip the FP32-mixed measurement")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-closure", action="store_true",
                    help="skip the learned-closure measurement (coarse 96x64x64 grid)")
    # about 10 s of the OpenMP oracle on 16 cores at 768x384x384
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--copy-probe", action="store_true",
                    help="time a plain device copy of 1 GiB (bandwidth ceiling and the "
                         "scale factor of the FETCH_SIZE/WRITE_SIZE counters) and exit")
    ap.add_argument("--closure-kernels", action="store_true",
                    help="run the learned closure's kernels through their own entry points "
                         "(filter, samples, training, both closure stages; no pll_step) on the "
                         "closure workflow's grids, print their HIP-event times and exit")
    ap.add_argument("--closure-only", action="store_true",
                    help="run only the closure measurement (the coarse 96x64x64 step with and "
                         "without the closure, and the training throughput), print it and exit")
    ap.add_argument("--stat-probe", action="store_true",
                    help="time the run-time statistics accumulation (all 11 sums) on the "
                         "workload grid in both precisions and exit")
    return ap.parse_args()


def closure_bench(steps=50, warmup=5):
    """The learned closure on the coarse grid of the Re_tau~395 closure
    workflow (96 x 64 x 64, 2pi x pi x 2, stretch 1.5, re = 6875, dt = 0.008,
    FP64): the step with and without a closure attached (its network
    evaluated once per cell for each RK stage), and Adam steps of the network's
    MFMA training kernels on one snapshot's worth of samples (393,216)."""
    import torch  # device buffers for the samples; imported before libpll.so
    import numpy as np

    import powerllel_channel_sycl_amd as pll
    nx, ny, nz, re = 96, 64, 64, 6875.0
    dzf = pll.mesh_by_func(1, 0, 1.5, nz, 2.0)
    utau = 395.0 / re
    cl = pll.Closure(seed=7)
    cl.set_scales(1.0 / (re * utau * utau), utau * re / 100.0, 1.0 / re)
    out = {"workload": "closure coarse 96x64x64 (Re_tau~395), FP64 step"}
    for name in ("plain", "closure"):
        with pll.Channel(nx, ny, nz, 2 * np.pi, np.pi, 2.0, re, 0.008, dzf) as ch:
            ch.initFlow("log", noise_intensity=0.1)
            if name == "closure":
                ch.attach_closure(cl)
            ch.step(warmup)
            ch.synchronize()
            t0 = time.perf_counter()
            ch.step(steps)
            ch.synchronize()
            ms = (time.perf_counter() - t0) / steps * 1e3
            ch.timing(True)
            ch.timing_reset()
            ch.step(10)
            ch.synchronize()
            t = ch.timing_read()
            stage = t["rk_closure" if name == "closure" else "rk1"]
            out[name] = {"ms_per_step": round(ms, 4),
                         "stage_us": round(stage[0] / max(stage[1], 1) * 1e3, 1)}
            if name == "closure":
                ch.attach_closure(None)
    n = nx * ny * nz
    g = torch.Generator(device="cuda").manual_seed(1)
    x = torch.randn((n, 10), device="cuda", generator=g)
    y = torch.rand((n,), device="cuda", generator=g)
    for _ in range(3):
        cl.train_step(x.data_ptr(), y.data_ptr(), n, 1e-3)
    cl.synchronize()
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        cl.train_step(x.data_ptr(), y.data_ptr(), n, 1e-3)
    cl.synchronize()
    sec = (time.perf_counter() - t0) / reps
    out["training"] = {"samples_per_adam_step": n, "us_per_adam_step": round(sec * 1e6, 1),
                       "samples_per_s": round(n / sec), "mfma_tflops": round(n * CLOSURE_TRAIN_FLOP / sec / 1e12, 1)}
    cl.close()
    return out


def closure_kernels_probe(reps=5):
    """The closure's kernels through their entry points, for per-kernel
    profiles: a fine FP32-mixed 384x256x192 channel filtered onto the coarse
    FP64 96x64x64 one with the training samples of one snapshot
    (pll_closure_samples: k_filter, k_samples), Adam steps on them
    (pll_closure_train_step: k_mlp, k_grad_reduce, k_adam), and the two RK
    stages of the coarse grid with the closure attached (pll_time_int_vel_rk1
    / _rk2: k_nut, k_rk_les) with the boundary updates between them."""
    import torch
    import numpy as np

    import powerllel_channel_sycl_amd as pll
    re_b, re_tau = 6875.0, 395.0
    utau = re_tau / re_b
    fine = pll.Channel(384, 256, 192, 2 * np.pi, np.pi, 2.0, re_b, 0.004,
                       pll.mesh_by_func(1, 0, 1.5, 192, 2.0), precision="fp32_mixed")
    coarse = pll.Channel(96, 64, 64, 2 * np.pi, np.pi, 2.0, re_b, 0.008,
                         pll.mesh_by_func(1, 0, 1.5, 64, 2.0))
    cl = pll.Closure(seed=7)
    cl.set_scales(1.0 / (re_b * utau * utau), utau * re_b / 100.0, 1.0 / re_b)
    n = 96 * 64 * 64
    x = torch.empty((n, 10), dtype=torch.float32, device="cuda:0")
    y = torch.empty((n,), dtype=torch.float32, device="cuda:0")
    out = {"workload": "closure kernels: fine 384x256x192 FP32-mixed, coarse 96x64x64 FP64"}
    try:
        fine.initFlow("log", noise_intensity=0.1)
        fine.updateBoundVel(0)
        coarse.timing(True)
        fine.timing(True)
        t0 = time.perf_counter()
        for _ in range(reps):
            cl.samples(fine, coarse, x.data_ptr(), y.data_ptr())
        coarse.synchronize()
        out["samples_us"] = round((time.perf_counter() - t0) / reps * 1e6, 1)
        t0 = time.perf_counter()
        for _ in range(reps):
            cl.train_step(x.data_ptr(), y.data_ptr(), n, 1e-3)
        cl.synchronize()
        out["adam_step_us"] = round((time.perf_counter() - t0) / reps * 1e6, 1)
        coarse.updateBoundVel(0)
        coarse.attach_closure(cl)
        coarse.timing_reset()
        for _ in range(reps):
            coarse.timeIntVelRK1()
            coarse.updateBoundVel(1)
            coarse.timeIntVelRK2()
            coarse.updateBoundVel(0)
        coarse.synchronize()
        t = coarse.timing_read()
        st = t["rk_closure"]
        out["closure_stage_us"] = round(st[0] / max(st[1], 1) * 1e3, 1)
        coarse.attach_closure(None)
    finally:
        cl.close()
        coarse.close()
        fine.close()
    print(json.dumps(out))


def copy_probe():
    """Plain device-to-device copy of 2^27 doubles (1 GiB read + 1 GiB written)."""
    import torch
    n = 1 << 27
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / reps
    gbs = 2 * n * 8 / sec / 1e9
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8,
                      "copy_GBps": round(gbs, 1), "copy_frac_of_peak": round(gbs / HBM_PEAK_GBS, 4)}))


# Statistics accumulation per point: u, v, w, p read once (the neighbours
# u(i-1), v(j-1) and w(k-1) come from cache), each of the 11 FP64 sums read
# and written.
STAT_BYTES = {"fp64": 4 * 8 + 11 * 16, "fp32_mixed": 4 * 4 + 11 * 16}


def stat_probe(wl, reps=10):
    """Wall-clock timing of pll_stat_accumulate (one launch each) on the
    workload grid, after two untimed calls, in both precisions."""
    out = {"workload": wl}
    for precision in ("fp64", "fp32_mixed"):
        ch, npts = build_channel(wl, precision)
        with ch:
            ch.stat_enable((1,) * 11)
            for _ in range(2):
                ch.stat_accumulate()
            ch.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                ch.stat_accumulate()
            ch.synchronize()
            us = (time.perf_counter() - t0) / reps * 1e6
            gbs = STAT_BYTES[precision] * npts / (us * 1e-6) / 1e9
            out[precision] = {"us_per_accumulate": round(us, 1),
                              "bytes_per_point": STAT_BYTES[precision],
                              "GBps": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    print(json.dumps(out))


def build_channel(wl, precision):
    import powerllel_channel_sycl_amd as pll
    nx, ny, nz, lx, ly, lz, r, re, dt, prof = WORKLOADS[wl]
    dzf = pll.mesh_by_func(1, 0, r, nz, lz)
    ch = pll.Channel(nx, ny, nz, lx, ly, lz, re, dt, dzf, precision=precision)
    ch.initFlow(prof, noise_intensity=0.1, seeds=(42, 43, 44))
    ch.updateBoundVel(0)
    ch.updateBoundP()
    return ch, nx * ny * nz


def cpu_baseline(wl, steps):
    """FP64 oracle on this host: `steps` RK2 steps of the full workload grid."""
    from oracle import pll_oracle as po
    nx, ny, nz, lx, ly, lz, r, re, dt, prof = WORKLOADS[wl]
    dzf = po.mesh_by_func(1, 0, r, nz, lz)
    o = po.Oracle(nx, ny, nz, lx, ly, lz, re, dt, dzf)
    u, v, w, u1, v1, w1, p = (o.field() for _ in range(7))
    o.init_flow(u, v, w, prof, noise_intensity=0.1)
    o.update_bound_vel(u, v, w)
    o.update_bound_p(p)
    o.step(1, u, v, w, u1, v1, w1, p)  # untimed: pages the fields in
    t0 = time.perf_counter()
    o.step(steps, u, v, w, u1, v1, w1, p)
    sec = time.perf_counter() - t0
    threads = int(po.lib().orc_num_threads())
    return {
        "value": nx * ny * nz * steps / sec / 1e6,
        "unit": "Mpoint-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} FP64 RK2 step(s) of the full {nx}x{ny}x{nz} grid after one untimed "
                  f"step, {sec:.2f} s, OpenMP oracle (the reference's SYCL build cannot be "
                  f"built without a SYCL compiler and oneMKL)",
        "cpu_model": cpu_model(),
        "omp_threads": threads,
        "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS"),
        "nproc": len(os.sched_getaffinity(0)),
        "cpu_count": os.cpu_count(),
    }


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def measure(wl, precision, steps, warmup):
    """Time `steps` RK2 steps after `warmup`, then re-run them under the per-kernel timers."""
    ch, npts = build_channel(wl, precision)
    print(f"# workload {wl} precision {precision}: {npts} points, "
          f"{ch.device_bytes() / 1e9:.2f} GB on device", flush=True)
    ch.step(warmup)
    ch.synchronize()
    windows = []
    for _ in range(WINDOWS):
        t0 = time.perf_counter()
        ch.step(steps)
        ch.synchronize()
        windows.append(time.perf_counter() - t0)
    sec = sorted(windows)[len(windows) // 2]
    ms_step = 1e3 * sec / steps
    print(f"# {precision}: {WINDOWS} windows of {steps} steps, "
          f"{', '.join(f'{1e3 * w / steps:.3f}' for w in windows)} ms/step; median "
          f"{ms_step:.3f}", flush=True)
    ch.timing(True)
    ch.timing_reset()
    ch.step(steps)
    ch.synchronize()
    times = ch.timing_read()
    ch.timing(False)
    mean_u = ch.Mean("u")
    div = ch.calcMaxDiv()[0]
    ch.close()
    scale = 1.0 if precision == "fp64" else 0.5
    table = BYTES_PER_POINT[precision]
    kernels = {}
    total = sum(ms for ms, n in times.values() if n)
    for name, (ms, n) in times.items():
        if n == 0:
            continue
        avg_ms = ms / n
        entry = {"avg_us": round(avg_ms * 1e3, 2), "launches": n, "share": round(ms / total, 4)}
        bpp = table.get(name)
        if bpp is not None:
            gbs = bpp * npts / (avg_ms * 1e-3) / 1e9
            entry["GBps"] = round(gbs, 1)
            entry["frac"] = round(gbs / HBM_PEAK_GBS, 4)
        kernels[name] = entry
    dom = max((k for k in kernels if "GBps" in kernels[k]), key=lambda k: times[k][0])
    d_avg_s = times[dom][0] / times[dom][1] * 1e-3
    achieved = table[dom] * npts / d_avg_s / 1e9
    measured = MEASURED_HBM_B_PER_POINT.get(precision, {}).get(dom)
    traffic = (round(measured * npts) if measured is not None and wl == DEFAULT_WORKLOAD
               else None)
    step_gbs = STEP_BYTES_FP64 * scale * npts / (ms_step * 1e-3) / 1e9
    return {
        "value": round(npts * steps / sec / 1e6, 2),
        "ms_per_step": round(ms_step, 4),
        "window_ms_per_step": [round(1e3 * w / steps, 4) for w in windows],
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "step_achieved": round(step_gbs, 1),
                     "step_frac": round(step_gbs / HBM_PEAK_GBS, 4),
                     "bytes_per_point_step": STEP_BYTES_FP64 * scale},
        "kernels": kernels,
        "check": {"bulk_u": mean_u, "max_div": div},
    }


def main():
    args = parse()
    if args.copy_probe:
        copy_probe()
        return
    if args.closure_kernels:
        closure_kernels_probe()
        return
    if args.closure_only:
        print(json.dumps({"closure": closure_bench()}), flush=True)
        return
    if args.stat_probe:
        stat_probe(args.workload)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus must be 1 "
                                   "(sharding across GPUs is designed, not built)"}))
        sys.exit(2)
    if not args.no_closure:
        # torch (device buffers of the closure's training samples) brings its
        # own HIP runtime, which must be loaded before libpll.so binds one
        import torch  # noqa: F401
    main64 = measure(args.workload, "fp64", args.steps, args.warmup)
    result = {
        "metric": "Mpoint-steps/s per RK2 step (FP64 vs FP32-mixed) + % HBM roofline, 1 GPU",
        "value": main64["value"],
        "unit": "Mpoint-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main64["ms_per_step"],
        "window_ms_per_step": main64["window_ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded log profile + 10% uniform noise)",
        "config": {"workload": args.workload, "grid": list(WORKLOADS[args.workload][:3]),
                   "precision": "fp64", "parallelism": "single GPU"},
        "roofline": main64["roofline"],
        "kernels": main64["kernels"],
        "check": main64["check"],
    }
    if not args.fp64_only:
        mixed = measure(args.workload, "fp32_mixed", args.steps, args.warmup)
        result["fp32_mixed"] = {
            "value": mixed["value"], "ms_per_step": mixed["ms_per_step"],
            "window_ms_per_step": mixed["window_ms_per_step"],
            "dtype": "f32 fields/stencils/FFTs, f64 tridiagonal arithmetic and bulk reduction",
            "speedup_vs_fp64": round(mixed["value"] / main64["value"], 3),
            "roofline": mixed["roofline"], "kernels": mixed["kernels"], "check": mixed["check"],
        }
    if not args.no_closure:
        result["closure"] = closure_bench()
    if not args.no_cpu_baseline:
        print("# cpu baseline ...", flush=True)
        result["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_steps)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
