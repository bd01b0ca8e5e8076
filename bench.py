"""Kirchhoff pair throughput on one MI355X.

Workload (one step): the Gaussian-source case, 1,048,576 rays x 512^2
pixels at 1000 mm, 10 keV (2.75e11 ray-pixel pairs), solved from a ray set
resident in HBM: FP64 centroid + prep, FP32 phase-reduced pair kernel, FP64
merge and base phase, result copied to the host.  The same solve with the
plain FP64 pair kernel runs beside it for the speed ratio and the error
(complex rel-L2 of FP32-reduced vs FP64 over all pixels).  The CPU baseline
is the oracle (the reference's f64 Kahan loop restated in C, OpenMP over
pixels) timed on a strided pixel sample of the same workload.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "xrt-rs_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

F_PAIR = 51                 # FP operations per pair (FMA = 2, transcendental = 1)
PEAK_FP32 = 157.3           # TFLOP/s, MI355X FP32 vector (packed) datasheet
PEAK_FP32_NOPACK = 78.65    # the same without packed FP32 (what scalar FP32 code can issue)
PEAK_FP64 = 78.6            # TFLOP/s, MI355X FP64 vector


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def device_sync():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipDeviceSynchronize()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--rays", type=int, default=1048576)
    ap.add_argument("--side", type=int, default=512)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the Kirchhoff engine has no multi-GPU path "
              "in this build (a pixel or ray shard per GPU is described in DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    import xrtk
    from xrtk import workloads as W

    ctx = xrtk.GpuContext.shared()
    if ctx is None:
        log("no gfx950 device")
        sys.exit(1)
    log(f"device: {ctx.adapter_name()}")
    t0 = time.time()
    rays = W.gaussian_source(args.rays, seed=2)
    scr = W.screen(1000.0, 0.02, args.side)
    wave = xrtk.WaveParams(scr.center, scr.dx, scr.dz, scr.nx, scr.nz)
    nr, npx = args.rays, args.side * args.side
    pairs = float(nr) * float(npx)
    log(f"inputs built in {time.time() - t0:.1f}s: {nr} rays x {npx} pixels = {pairs:.3e} pairs")
    rs = xrtk.RaySet(ctx, rays)
    job = xrtk.Job(rs, 10000.0, wave)

    # ---- FP32-reduced (the product path)
    for i in range(args.warmup):
        xrtk.solve_jobs(ctx, [job], mode="f32r")
        log(f"warmup {i + 1}/{args.warmup}: pair {ctx.last_timing()[0]:.1f} ms")
    device_sync()
    pair_ms = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        (out32,) = xrtk.solve_jobs(ctx, [job], mode="f32r")
        pair_ms.append(ctx.last_timing()[1])
    device_sync()
    t_total = time.perf_counter() - t_start
    ms_step = 1e3 * t_total / args.steps
    value = pairs / (t_total / args.steps) / 1e9
    log(f"f32r: {ms_step:.1f} ms/step  {value:.1f} Gpair/s  (device graph {np.median(pair_ms):.1f} ms)")

    # per-launch duration of the pair kernel alone (HIP events on the
    # engine's stream), from one eager solve of the same instance
    xrtk.diffraction_integral(rays, wave, ctx=ctx, mode="f32r")
    k_ms = []
    for _ in range(3):
        xrtk.diffraction_integral(rays, wave, ctx=ctx, mode="f32r")
        k_ms.append(ctx.last_timing()[0])
    kernel_ms = float(np.median(k_ms))
    plan = ctx.last_plan()
    achieved = pairs * F_PAIR / (kernel_ms * 1e-3) / 1e12
    log(f"pair kernel: {kernel_ms:.1f} ms/launch  {achieved:.1f} TFLOP/s  plan {plan}")

    # ---- plain FP64 (speed ratio and error reference)
    fp64 = None
    if not args.no_fp64 and args.fp64_steps > 0:
        (out64,) = xrtk.solve_jobs(ctx, [job], mode="f64")
        t64 = []
        for i in range(args.fp64_steps):
            t = time.perf_counter()
            (out64,) = xrtk.solve_jobs(ctx, [job], mode="f64")
            t64.append(time.perf_counter() - t)
            log(f"f64 step {i + 1}: {t64[-1] * 1e3:.0f} ms")
        v64 = pairs / float(np.median(t64)) / 1e9
        err = W.rel_l2(out32, out64)
        imax, ip99 = W.intensity_error(out32, out64)
        fp64 = {"value": v64, "unit": "Gpair-evals/s", "ms_per_step": 1e3 * float(np.median(t64)),
                "speedup_f32r_over_f64": value / v64, "rel_l2_f32r_vs_f64": err,
                "intensity_err_max": imax, "intensity_err_p99": ip99,
                "roofline_frac_fp64": v64 * 1e9 * F_PAIR / 1e12 / PEAK_FP64}
        log(f"f64: {v64:.2f} Gpair/s, speedup {value / v64:.2f}x, rel-L2 {err:.3e}")

    # ---- CPU baseline (oracle, strided pixel sample)
    cpu = None
    if not args.no_cpu:
        import pyoracle
        threads = min(16, len(os.sched_getaffinity(0)))
        px, py, pz = xrtk.prepare_wave(wave)
        stride0 = max(1, npx // 8)
        t = time.perf_counter()
        pyoracle.diffraction_f64(rays, px, py, pz, stride=stride0, threads=threads)
        dt = time.perf_counter() - t
        n0 = (npx + stride0 - 1) // stride0
        per_px = dt / n0
        want = max(threads, int(args.cpu_seconds / per_px))
        stride = max(1, npx // want)
        nsamp = (npx + stride - 1) // stride
        log(f"cpu baseline: {threads} threads, {nsamp} pixels (stride {stride}), est {nsamp * per_px:.1f}s")
        t = time.perf_counter()
        ref = pyoracle.diffraction_f64(rays, px, py, pz, stride=stride, threads=threads)
        dt = time.perf_counter() - t
        cpu_rate = nsamp * nr / dt / 1e9
        cpu = {"value": cpu_rate, "unit": "Gpair-evals/s", "cores": threads, "kind": "port",
               "sample": f"{nsamp} pixels (every {stride}th of 512^2) x {nr} rays, "
                         f"{dt:.1f} s, f64 Kahan oracle, OpenMP over pixels"}
        cpu["rel_l2_f32r_vs_oracle_on_sample"] = W.rel_l2(out32[::stride], ref)
        log(f"cpu: {cpu_rate:.4f} Gpair/s; rel-L2 f32r vs oracle on sample "
            f"{cpu['rel_l2_f32r_vs_oracle_on_sample']:.3e}")
    rs.release()

    line = {
        "metric": "Gpair-evals/s FP32-reduced vs FP64 at relL2<=1e-4; served p50/p99 latency",
        "value": value, "unit": "Gpair-evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Gaussian source)",
        "config": {"workload": "gaussian-source Kirchhoff solve, 1,048,576 rays x 512^2 pixels "
                               "at 1000 mm, 10 keV, resident ray set",
                   "rays": nr, "pixels": npx, "pairs": pairs, "plan": plan},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32, "frac_vs_unpacked_fp32": achieved / PEAK_FP32_NOPACK,
                     "flop_per_pair": F_PAIR, "kernel_ms": kernel_ms, "traffic": None},
        "fp64": fp64,
        "cpu_baseline": cpu,
        "serving": None,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
