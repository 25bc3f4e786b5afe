"""Benchmark: PlanarCDI HIO iterations per second at 2048 x 2048 on one GPU.

A step is one HIO iteration (beta 0.9) of the 2048^2 star-object workload
(beam stop 16x16, support 900x900) with the estimate, intensity and support
resident in HBM.  `value` is the FP32-bulk fused path (FP64 error sums); the
FP64 paths (plain = one kernel per reference operation, fused) are timed
beside it on the same input, and the CPU oracle on a bounded sample.

Prints one JSON line.  Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import ctypes
import importlib
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "HIO iter/s at 2048² (FP32 batched vs FP64 plain, 1 GPU); served p50/p99 ms"
N = 2048

# Algorithmic bytes per pixel (FP32: complex 8 B, real 4 B; FP64 doubles it).
#   column pass: read W, read sqrt(I), write W          = 2 b_c + b_r
#   row pass (HIO): read W, x, support; write x, W      = 4 b_c + b_r
# Whole HIO iteration per pixel as counted for the roofline of the fused
# design (row/column pair with on-chip 1-D transforms): 7 b_c + 2 b_r.
def pass_bytes(double):
    bc, br = (16, 8) if double else (8, 4)
    return {"col": 2 * bc + br, "row": 4 * bc + br, "iter": 7 * bc + 2 * br}


HBM_PEAK_GBS = 8000.0


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def setup(nd, data, double, path):
    cdt = np.complex128 if double else np.complex64
    est = nd.Complex2D(N, N, double=double, data=data["x0"].astype(cdt))
    p = nd.PlanarCDI(est)
    p.set_path(path)
    p.set_support(data["support"])
    p.set_intensity(data["intensity"])
    p.set_beam_stop(data["beam_stop"])
    p.set_algorithm(nd.HIO)
    return est, p


def time_iterations(nd, p, steps, warmup):
    p.iterate_async(warmup)
    nd.synchronize()
    t0 = time.perf_counter()
    p.iterate_async(steps)
    nd.synchronize()
    return time.perf_counter() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--fp64-steps", type=int, default=40)
    ap.add_argument("--cpu-sample", type=int, default=12, help="oracle HIO iterations timed on the host")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp32-only", action="store_true",
                    help="run only the FP32 fused iterations (for profiler runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="K",
                    help="instead of the benchmark, run K device copies of a 2048^2 complex64 field "
                         "(32 MiB read + 32 MiB written each) to calibrate profiler byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(independent reconstructions shard by job across GPUs)", file=sys.stderr)
        sys.exit(2)

    nd = importlib.import_module("nadia-coecxs_amd")
    synth = importlib.import_module("nadia-coecxs_amd.synth")
    if args.copy_calibration:
        a = nd.Complex2D(N, N, double=False)
        b = nd.Complex2D(N, N, double=False)
        for _ in range(args.copy_calibration):
            b.copy(a)
        nd.synchronize()
        print(json.dumps({"copy_calibration": args.copy_calibration, "bytes_read_per_copy": N * N * 8,
                          "bytes_written_per_copy": N * N * 8}))
        return
    t_gen = time.perf_counter()
    data = synth.planar_star(n=N)
    t_gen = time.perf_counter() - t_gen
    print("bench: inputs generated in %.1f s" % t_gen, file=sys.stderr, flush=True)

    # ---- FP32 fused, batched (the headline value) ----------------------
    # one job alone, then two independent 2048^2 jobs sharing every launch;
    # value = the better aggregate job-iterations per second
    est, p = setup(nd, data, False, nd.PATH_FUSED)
    dt = time_iterations(nd, p, args.steps, args.warmup)
    single = args.steps / dt
    err32 = p.get_error()
    print("bench: fp32 fused, 1 job: %.1f iter/s" % single, file=sys.stderr, flush=True)
    est2, p2 = setup(nd, data, False, nd.PATH_FUSED)
    batch = nd.Batch([p, p2])
    batch.iterate_async(args.warmup)
    nd.synchronize()
    t0 = time.perf_counter()
    batch.iterate_async(args.steps)
    nd.synchronize()
    dual = 2 * args.steps / (time.perf_counter() - t0)
    print("bench: fp32 fused, batch of 2 jobs: %.1f job-iter/s" % dual, file=sys.stderr, flush=True)
    del batch, p2, est2
    batch_jobs = 2 if dual > single else 1
    value = max(single, dual)
    # per-pass device time with HIP events on the library stream (one job)
    p.set_timing(True)
    p.iterate_async(min(args.steps, 200))
    nd.synchronize()
    ms, cnt = p.pass_times()
    p.set_timing(False)
    avg = {"col": ms[0] / max(cnt[0], 1), "row": ms[1] / max(cnt[1], 1)}
    b32 = pass_bytes(False)
    dom = max(avg, key=avg.get)
    bytes_launch = b32[dom] * N * N
    achieved = bytes_launch / (avg[dom] * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_hio2048_fp32.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get(dom)
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
        "kernel": "col_kernel<float,2048,COL_PLANAR>" if dom == "col" else "row_kernel<float,2048,ROW_UPDATE>",
        "bytes_per_launch": bytes_launch, "avg_launch_ms": round(avg[dom], 5),
        "col_pass_ms": round(avg["col"], 5), "row_pass_ms": round(avg["row"], 5),
        "tile_variants": list(p.get_tile_variants()),
        "iteration_achieved_GBs": round(b32["iter"] * N * N * single / 1e9, 1),
    }
    del p, est
    if args.fp32_only:
        print(json.dumps({"fp32_fused_iter_s": round(single, 2), "fp32_batched_job_iter_s": round(dual, 2),
                          "roofline": roofline}))
        return

    # ---- FP64 plain (reference data flow) and FP64 fused ----------------
    est64, p64 = setup(nd, data, True, nd.PATH_PLAIN)
    dt64 = time_iterations(nd, p64, args.fp64_steps, 3)
    plain64 = args.fp64_steps / dt64
    print("bench: fp64 plain %.1f iter/s" % plain64, file=sys.stderr, flush=True)
    p64.set_path(nd.PATH_FUSED)
    dtf = time_iterations(nd, p64, args.fp64_steps * 5, 5)
    fused64 = args.fp64_steps * 5 / dtf
    print("bench: fp64 fused %.1f iter/s" % fused64, file=sys.stderr, flush=True)
    del p64, est64

    # ---- CPU baseline: the oracle on the host, bounded sample -----------
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle  # CPU restatement of the reference (checker / baseline only)

        res = {}
        for double in (False, True):
            cdt = np.complex128 if double else np.complex64
            xo = data["x0"].astype(cdt)
            o = oracle.OracleCDI(xo, double=double)
            o.set_support(data["support"])
            o.set_intensity(data["intensity"])
            o.set_beam_stop(data["beam_stop"])
            o.set_algorithm(oracle.HIO)
            o.iterate()  # warm-up (temporaries, first touch)
            t0 = time.perf_counter()
            o.iterate_n(args.cpu_sample)
            res[double] = args.cpu_sample / (time.perf_counter() - t0)
            print("bench: cpu oracle %s %.3f iter/s" % ("fp64" if double else "fp32", res[double]),
                  file=sys.stderr, flush=True)
        cpu = {"value": round(res[False], 4), "unit": "iter/s", "cores": 1, "kind": "port",
               "sample": "%d HIO iterations at 2048^2 after 1 warm-up, FP32 arrays with FP64 error sums "
                         "(the reference's default precision), single thread, oracle's own FFT (no FFTW); "
                         "host CPU: %s" % (args.cpu_sample, cpu_model()),
               "fp64_value": round(res[True], 4)}

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "iter/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1000.0 / value, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "PlanarCDI HIO beta=0.9, 2048x2048 star-shaped complex object, "
                               "Poisson peak 1e8, 16x16 beam stop, 900x900 support, FP32 bulk + FP64 error sums",
                   "nx": N, "ny": N, "algorithm": "HIO", "beta": 0.9, "batch_jobs": batch_jobs},
        "fp32_single_job_iter_s": round(single, 2), "fp32_batch2_job_iter_s": round(dual, 2),
        "fp64_plain_iter_s": round(plain64, 2), "fp64_fused_iter_s": round(fused64, 2),
        "fp32_vs_fp64_plain": round(value / plain64, 2),
        "final_error_fp32": err32,
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
