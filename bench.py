"""Benchmark: PlanarCDI HIO iterations per second at 2048 x 2048 on one GPU.

A step is one HIO iteration (beta 0.9) of the 2048^2 star-object workload
(beam stop 16x16, support 900x900) with the estimate, intensity and support
resident in HBM.  `value` is the FP32-bulk fused path (FP64 error sums),
batched; the FP64 paths (plain = one kernel per reference operation, fused)
are timed beside it on the same input, and the CPU oracle on a bounded
sample.  Further legs (each can be skipped):
  batched64  64 independent RAAR jobs of sizes {256..1024}^2 (FP32), eager
             launches and graph replays
  served     a Poisson request stream through the continuous-batching loop
             (p50/p99 latency, aggregate iterations/s)
  accuracy   final-object error of the FP32 path against the FP64 path for
             the 2048^2 HIO + shrinkwrap + ER schedule, the 64-job batch and
             the 1024^2 Fresnel workload
  phase_diverse  8 Fresnel frames of 1024^2 merged by PhaseDiverseCDI
  next_rows  FresnelCDI ER with a TransmissionConstraint, HIO keeping the 4
             best estimates, the intensity autocorrelation

Prints one JSON line.  Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import dataclasses
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
#   row pass (HIO): read W, x, support; write x, W      = 4 b_c + 1
#     (a support of zeros and ones, as every workload here has, is read as
#     one byte per pixel)
# Whole HIO iteration per pixel as counted for the roofline of the fused
# design (row/column pair with on-chip 1-D transforms): 7 b_c + 2 b_r.
def pass_bytes(double):
    bc, br = (16, 8) if double else (8, 4)
    return {"col": 2 * bc + br, "row": 4 * bc + 1, "iter": 7 * bc + 2 * br}


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


def rel_l2(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _best_shift(a, b):
    """a rolled by the cyclic shift that maximises |cross-correlation| with b."""
    c = np.fft.ifft2(np.fft.fft2(b) * np.conj(np.fft.fft2(a)))
    k = np.unravel_index(int(np.argmax(np.abs(c))), c.shape)
    return np.roll(a, k, axis=(0, 1))


def _twins(a):
    """a and its twin image conj(a(-r)) (same far-field modulus)."""
    return (a, np.conj(np.roll(a[::-1, ::-1], 1, axis=(0, 1))))


def aligned_rel_l2(a, b, far_field=True, scale=False):
    """Relative L2 of a against b up to what the measurement cannot fix: a
    global phase (complex scale with scale=True, for a comparison with the
    true object, whose amplitude the intensity scaling changed), and for a
    far-field (PlanarCDI) reconstruction also a cyclic translation and the
    twin image; the best over those."""
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    best = np.inf
    for c in (_twins(a) if far_field else (a,)):
        if far_field:
            c = _best_shift(c, b)
        ip = np.vdot(c, b)
        if scale:
            c = c * (ip / max(np.vdot(c, c).real, 1e-300))
        elif abs(ip) > 0:
            c = c * (ip / abs(ip))
        best = min(best, rel_l2(c, b))
    return float(best)


def leg_accuracy_planar(nd, synth, data, hio=1000, every=20, er=100):
    """The 2048^2 schedule (HIO beta 0.9 with shrinkwrap every `every`
    iterations, sigma 3 -> 1.5, threshold 0.1, then ER) in FP32 and FP64
    from the same start; FP64 once more from a start perturbed by 1e-15
    (relative) to show how far rounding alone moves this trajectory."""
    ncalls = hio // every
    out = {}
    finals = {}
    for tag, double, eps in (("fp32", False, 0.0), ("fp64", True, 0.0), ("fp64_perturbed", True, 1e-15)):
        d = dict(data)
        if eps:
            rng = np.random.default_rng(99)
            d["x0"] = data["x0"] * (1 + eps * rng.standard_normal(data["x0"].shape))
        est, p = setup(nd, d, double, nd.PATH_FUSED)
        # tile-variant choices (HIO, and ER, whose work-field layout has its
        # own) and the first transforms outside the clock
        p.iterate_async(1)
        p.set_algorithm(nd.ER)
        p.iterate_async(1)
        p.set_algorithm(nd.HIO)
        nd.synchronize()
        est.set(d["x0"].astype(np.complex128 if double else np.complex64))
        t0 = time.perf_counter()
        for k in range(ncalls):
            p.iterate_async(every)
            p.apply_shrinkwrap(synth.shrinkwrap_sigma(k, ncalls), 0.1)
        p.set_algorithm(nd.ER)
        p.iterate_async(er)
        nd.synchronize()
        out[tag + "_s"] = round(time.perf_counter() - t0, 4)
        out[tag + "_final_error"] = p.get_error()
        finals[tag] = est.get()
        del p, est
    out["fp32_vs_fp64_rel_l2"] = rel_l2(finals["fp32"], finals["fp64"])
    out["fp64_perturbed_vs_fp64_rel_l2"] = rel_l2(finals["fp64_perturbed"], finals["fp64"])
    # up to global phase, translation and twin image
    out["fp32_vs_fp64_aligned_rel_l2"] = aligned_rel_l2(finals["fp32"], finals["fp64"])
    out["fp64_perturbed_vs_fp64_aligned_rel_l2"] = aligned_rel_l2(finals["fp64_perturbed"], finals["fp64"])
    # against the true object (complex scale, translation, twin image removed)
    for tag in finals:
        out[tag + "_vs_true_object_rel_l2"] = aligned_rel_l2(finals[tag], data["obj"], scale=True)
    out["schedule"] = "%d HIO (beta 0.9) with shrinkwrap every %d (sigma 3->1.5, threshold 0.1), then %d ER" % (
        hio, every, er)
    return out


def leg_batched(nd, synth, iters, graph_steps):
    """64 independent RAAR (beta 0.85) jobs of sizes drawn from
    {256,384,512,768,1024}^2: FP32 eager launches, FP32 graph replays
    (`graph_steps` iterations per graph) and FP64 eager, each from the same
    start for `iters` iterations in total; final-object error FP32 vs FP64."""
    jobs = synth.batched_jobs()

    def make(double):
        cdt = np.complex128 if double else np.complex64
        out = []
        for d in jobs:
            n = d["n"]
            est = nd.Complex2D(n, n, double=double, data=d["x0"].astype(cdt))
            g = nd.PlanarCDI(est)
            g.set_support(d["support"])
            g.set_intensity(d["intensity"])
            g.set_relaxation_parameter(0.85)
            g.set_algorithm(nd.RAAR)
            out.append((est, g))
        return out

    px = sum(d["n"] ** 2 for d in jobs)
    warm = graph_steps
    timed = iters - warm
    res = {"jobs": len(jobs), "pixels_per_sweep": px, "iterations": iters, "timed_iterations": timed,
           "graph_steps": graph_steps,
           "sizes": {str(n): sum(1 for d in jobs if d["n"] == n) for n in sorted({d["n"] for d in jobs})}}
    finals = {}
    for tag, double, graph in (("fp32_eager", False, False), ("fp32_graph", False, True), ("fp64_eager", True, False),
                               ("fp64_perturbed", True, False)):
        js = make(double)
        if tag == "fp64_perturbed":
            # the same start perturbed by 1e-15 (relative): how far rounding
            # alone moves these trajectories
            for (est, _), dj in zip(js, jobs):
                rng = np.random.default_rng(99)
                est.set(dj["x0"] * (1 + 1e-15 * rng.standard_normal(dj["x0"].shape)))
        b = nd.Batch([g for _, g in js])
        if graph:
            b.set_graph(graph_steps)
        b.iterate_async(warm)
        nd.synchronize()
        t0 = time.perf_counter()
        b.iterate_async(timed)
        nd.synchronize()
        dt = time.perf_counter() - t0
        bpp = pass_bytes(double)["iter"]
        res[tag] = {"job_iter_s": round(len(jobs) * timed / dt, 1), "ms_per_sweep": round(dt / timed * 1e3, 4),
                    "algorithmic_GBs": round(bpp * px * timed / dt / 1e9, 1)}
        if graph:
            res[tag]["graph_replays"] = b.graph_replays()
        finals[tag] = [est.get() for est, _ in js]
        del b, js
    res["graph_bit_identical_to_eager"] = all(np.array_equal(a, b) for a, b in
                                              zip(finals["fp32_eager"], finals["fp32_graph"]))
    def stats(errs):
        return {"median": float(np.median(errs)), "max": float(np.max(errs)),
                "jobs_le_1e-4": int(sum(e <= 1e-4 for e in errs))}

    for key, other in (("fp32_vs_fp64", "fp32_graph"), ("fp64_perturbed_vs_fp64", "fp64_perturbed")):
        res[key + "_rel_l2"] = stats([rel_l2(a, b) for a, b in zip(finals[other], finals["fp64_eager"])])
        res[key + "_aligned_rel_l2"] = stats([aligned_rel_l2(a, b) for a, b in
                                              zip(finals[other], finals["fp64_eager"])])
    for tag, key in (("fp32_graph", "fp32"), ("fp64_eager", "fp64"), ("fp64_perturbed", "fp64_perturbed")):
        res[key + "_vs_true_object_rel_l2"] = stats([aligned_rel_l2(a, d["obj"], scale=True)
                                                     for a, d in zip(finals[tag], jobs)])
    return res


def leg_fresnel(nd, synth, n=1024, wf_iters=50, er_iters=300):
    """Fresnel workload: white field recovered by FresnelCDI_WF iterations on
    its intensity, then FresnelCDI ER on the object intensity, FP32 and FP64;
    iterations/s and final errors FP32 vs FP64 (the object stage of both
    precisions starts from the FP64-recovered white field, so its error is
    that of the ER iterations alone)."""
    d = synth.fresnel_config(n=n)
    out = {"n": n, "wf_iterations": wf_iters, "er_iterations": er_iters}
    fin = {}
    wfs = {}
    for tag, double in (("fp64", True), ("fp32", False)):
        wfe = nd.Complex2D(n, n, double=double)
        w = nd.FresnelCDI_WF(wfe, d["wavelength"], d["z12"], d["z23"], d["pixel"])
        w.set_support(d["pupil_diameter"])
        w.set_intensity(d["white_field_intensity"])
        # one iteration first (tile-variant choice, twiddle tables), then the
        # timed run from the initial estimate
        w.initialise_estimate(0)
        w.iterate_async(1)
        nd.synchronize()
        w.initialise_estimate(0)
        t0 = time.perf_counter()
        w.iterate_async(wf_iters)
        nd.synchronize()
        t_wf = time.perf_counter() - t0
        wfs[tag] = wfe.get()
        wf = wfs["fp64"]
        est = nd.Complex2D(n, n, double=double)
        f = nd.FresnelCDI(est, wf, d["wavelength"], d["zfd"], d["zfs"], d["pixel"], 1.0)
        f.set_support(d["support"])
        f.set_intensity(d["intensity"])
        f.initialise_estimate(0)
        f.iterate_async(1)
        nd.synchronize()
        f.initialise_estimate(0)
        t0 = time.perf_counter()
        f.iterate_async(er_iters)
        nd.synchronize()
        t_er = time.perf_counter() - t0
        out[tag] = {"wf_iter_s": round(wf_iters / t_wf, 1), "er_iter_s": round(er_iters / t_er, 1),
                    "final_error": f.get_error()}
        fin[tag] = (est.get(), f.get_transmission_function())
        del f, est, w, wfe
    out["white_field_fp32_vs_fp64_rel_l2"] = rel_l2(wfs["fp32"], wfs["fp64"])
    out["esw_fp32_vs_fp64_rel_l2"] = rel_l2(fin["fp32"][0], fin["fp64"][0])
    out["transmission_fp32_vs_fp64_rel_l2"] = rel_l2(fin["fp32"][1], fin["fp64"][1])
    # against the truth: the white field up to its global phase (|L_d|^2
    # leaves it free), the transmission inside the support as it is (the
    # white field fixes its phase)
    sup = d["support"] > 0
    for tag in ("fp32", "fp64"):
        out["white_field_%s_vs_true_rel_l2" % tag] = aligned_rel_l2(wfs[tag], d["white_field"], far_field=False)
        out["transmission_%s_vs_true_rel_l2" % tag] = rel_l2(fin[tag][1][sup], d["transmission"][sup])
    return out


def leg_reference_example(nd, synth):
    """The reference's own example: the measured 1024^2 pattern of
    NADIA/examples/image_files/planar_data.tif through the schedule of
    NADIA/examples/PlanarCDI_example.c (2 cycles of 50 ER, 100 HIO at
    relaxation 0.5, 50 ER, shrinkwrap every 50, 4 best estimates kept on
    the device), one iterate() call and one get_error() per iteration as the
    example's loop does (it prints every error), FP32 and FP64: wall time of
    the whole schedule and the final errors."""
    d = synth.reference_planar_example(os.path.join(ROOT, "tests", "golden", "planar_data.npz"))
    n = d["intensity"].shape[0]
    out = {"n": n, "input": "NADIA/examples/image_files/planar_data.tif",
           "schedule": "PlanarCDI_example.c: 2 x (50 ER, 100 HIO beta 0.5, 50 ER), shrinkwrap every 50, n_best 4"}
    fin = {}
    for tag, double in (("fp32", False), ("fp64", True)):
        est = nd.Complex2D(n, n, double=double)
        p = nd.PlanarCDI(est, n_best=4)
        p.set_support(d["support"])
        p.set_intensity(d["intensity"])
        p.initialise_estimate(0)
        algs = {"ER": nd.ER, "HIO": nd.HIO}
        t0 = time.perf_counter()
        errs = synth.run_reference_example(
            p, alg=lambda a: p.set_algorithm(algs[a]), iterate=lambda: float(p.iterate(1)[0]),
            shrinkwrap=p.apply_shrinkwrap, set_beta=p.set_relaxation_parameter)
        nd.synchronize()
        dt = time.perf_counter() - t0
        best, best_err = p.get_best_result(0)
        out[tag] = {"iterations": len(errs), "seconds": round(dt, 3), "iter_s": round(len(errs) / dt, 1),
                    "final_error": errs[-1], "best_error": best_err}
        fin[tag] = est.get()
        del p, est
    out["fp32_vs_fp64_rel_l2"] = rel_l2(fin["fp32"], fin["fp64"])
    return out


def leg_phase_diverse(nd, synth, n=1024, frames=8, big_iters=10):
    """PhaseDiverseCDI: `frames` Fresnel frames of n^2 (the 1024^2 Fresnel
    workload, shifted by 16 px per frame along x) merged in parallel mode,
    one ER iteration per frame per big iteration, FP32 and FP64."""
    d = synth.fresnel_config(n=n)
    out = {"n": n, "frames": frames, "big_iterations": big_iters, "mode": "parallel, beta 0.9, gamma 1"}
    for tag, double in (("fp32", False), ("fp64", True)):
        cdt = np.complex128 if double else np.complex64
        keep = []
        pd = nd.PhaseDiverseCDI(beta=0.9, gamma=1.0, parallel=True)
        for k in range(frames):
            est = nd.Complex2D(n, n, double=double)
            f = nd.FresnelCDI(est, d["white_field"].astype(cdt), d["wavelength"], d["zfd"], d["zfs"], d["pixel"], 1.0)
            f.set_support(d["support"])
            f.set_intensity(d["intensity"])
            f.initialise_estimate(k)
            pd.add_new_position(f, 16.0 * k, 0.0)
            keep.append((est, f))
        pd.initialise_estimate()
        pd.iterate()  # warm-up: weights, tile variants
        nd.synchronize()
        t0 = time.perf_counter()
        for _ in range(big_iters):
            e = pd.iterate()
        nd.synchronize()
        dt = time.perf_counter() - t0
        out[tag] = {"big_iter_s": round(big_iters / dt, 2), "frame_iter_s": round(frames * big_iters / dt, 1),
                    "object": list(pd.size()), "mean_frame_error": float(np.mean(e))}
        del pd, keep
    return out


def leg_next_rows(nd, synth, data2048):
    """The components either side of the loop: FresnelCDI 1024^2 ER with a
    TransmissionConstraint (one complex-constraint region over the support,
    unity magnitude and charge flipping on: the unfused path), PlanarCDI
    2048^2 HIO FP32 keeping the 4 best estimates on the device, and the
    intensity autocorrelation at 2048^2."""
    out = {}
    d = synth.fresnel_config(n=1024)
    n = 1024
    rng = np.random.default_rng(5)
    T0 = np.exp((0.3 + 1j) * -rng.uniform(0.1, 0.4, (n, n)))
    for tag, double in (("fp32", False), ("fp64", True)):
        cdt = np.complex128 if double else np.complex64
        est = nd.Complex2D(n, n, double=double)
        f = nd.FresnelCDI(est, d["white_field"].astype(cdt), d["wavelength"], d["zfd"], d["zfs"], d["pixel"], 1.0)
        f.set_support(d["support"])
        f.set_intensity(d["intensity"])
        tc = nd.TransmissionConstraint(n, n)
        tc.add_complex_constraint(d["support"], 0.5, 0.5)
        f.set_complex_constraint(tc)
        f.set_transmission_function(T0)
        f.iterate_async(2)
        nd.synchronize()
        f.set_transmission_function(T0)
        iters = 100
        t0 = time.perf_counter()
        f.iterate_async(iters)
        nd.synchronize()
        out["fresnel_constraint_1024_" + tag + "_iter_s"] = round(iters / (time.perf_counter() - t0), 1)
        del f, tc, est
    est = nd.Complex2D(N, N, double=False, data=data2048["x0"].astype(np.complex64))
    p = nd.PlanarCDI(est, n_best=4)
    p.set_support(data2048["support"])
    p.set_intensity(data2048["intensity"])
    p.set_beam_stop(data2048["beam_stop"])
    p.set_algorithm(nd.HIO)
    p.iterate_async(20)
    nd.synchronize()
    t0 = time.perf_counter()
    p.iterate_async(300)
    nd.synchronize()
    out["hio_2048_fp32_n_best4_iter_s"] = round(300 / (time.perf_counter() - t0), 1)
    p.get_intensity_autocorrelation()
    t0 = time.perf_counter()
    for _ in range(10):
        p.get_intensity_autocorrelation()
    out["autocorrelation_2048_fp32_ms"] = round((time.perf_counter() - t0) * 100, 3)
    del p, est
    # sizes without compiled passes (any-length transforms, unfused
    # iteration) beside the native 1024^2: 1000 = 8*5^3 (direct plan), 1030 =
    # 2*5*103 (Bluestein, transforms of length 4096)
    for n in (1024, 1000, 1030):
        d = synth.planar_block(n=n, block=3 * n // 8, support=7 * n // 16)
        est = nd.Complex2D(n, n, double=False, data=d["x0"].astype(np.complex64))
        p = nd.PlanarCDI(est)
        p.set_support(d["support"])
        p.set_intensity(d["intensity"])
        p.set_algorithm(nd.HIO)
        p.iterate_async(10)
        nd.synchronize()
        t0 = time.perf_counter()
        p.iterate_async(200)
        nd.synchronize()
        out["hio_%d_fp32_iter_s" % n] = round(200 / (time.perf_counter() - t0), 1)
        del p, est
    return out


def leg_served(nd, synth, count, rate, budget):
    """Poisson request stream through the continuous-batching loop (FP32
    bulk, FP64 reductions), admission by an HBM budget."""
    serve = importlib.import_module("nadia-coecxs_amd.serve")
    reqs = serve.poisson_stream(count=count, rate=rate)
    sizes = sorted({r.n for r in reqs})
    t0 = time.perf_counter()
    inputs = serve.input_pool(synth, sizes, per_size=2)
    t_gen = time.perf_counter() - t0
    be = serve.DeviceBackend(nd, inputs)
    # one short job per size first: tile-variant choice and twiddle tables
    # happen outside the measured stream
    serve.Server(be, hbm_budget=budget).run(
        [serve.Request(-1 - i, 0.0, n, serve.HIO, 0.9, 20, False, 0) for i, n in enumerate(sizes)])
    srv = serve.Server(be, hbm_budget=budget)
    done = srv.run(reqs)
    s = serve.Server.summary(done, srv.elapsed)
    s.update({"rate_req_s": rate, "hbm_budget_GB": budget / 1e9, "peak_active": srv.peak_active,
              "peak_bytes_GB": round(srv.peak_bytes / 1e9, 3), "ticks": srv.ticks,
              "input_pool_s": round(t_gen, 1),
              "p50_ms": round(s["p50_ms"], 3), "p99_ms": round(s["p99_ms"], 3),
              "aggregate_iter_s": round(s["aggregate_iter_s"], 1)})
    by = {}
    for c in done:
        by.setdefault(c.n, []).append(c.latency * 1e3)
    s["p50_ms_by_size"] = {str(n): round(float(np.percentile(v, 50)), 3) for n, v in sorted(by.items())}
    s["p99_ms_by_size"] = {str(n): round(float(np.percentile(v, 99)), 3) for n, v in sorted(by.items())}
    s["capacity"] = leg_capacity(nd, serve, reqs, be, budget)
    return s


def leg_capacity(nd, serve, reqs, be, budget, count=400):
    """Capacity: the first `count` requests of the served mix arriving at
    once (the queue drains at the device's pace); sustained iterations/s,
    jobs/s and pixel-iterations/s."""
    burst = [dataclasses.replace(r, arrival=0.0) for r in reqs[:count]]
    srv = serve.Server(be, hbm_budget=budget)
    done = srv.run(burst)
    cap = serve.Server.summary(done, srv.elapsed)
    px = sum(c.iterations * c.n * c.n for c in done)
    return {"requests": len(burst), "elapsed_s": round(srv.elapsed, 3),
            "aggregate_iter_s": round(cap["aggregate_iter_s"], 1),
            "pixel_iter_per_s": round(px / srv.elapsed, 1),
            "jobs_per_s": round(len(burst) / srv.elapsed, 1), "peak_active": srv.peak_active,
            "ticks": srv.ticks}


def profile_leg(nd, synth, args):
    """One leg alone (for profiler runs): its JSON on stdout."""
    leg = args.profile
    if leg == "batched":
        out = leg_batched(nd, synth, args.batched_iters, args.graph_steps)
    elif leg == "fresnel":
        out = leg_fresnel(nd, synth)
    elif leg == "phase_diverse":
        out = leg_phase_diverse(nd, synth)
    elif leg == "next_rows":
        out = leg_next_rows(nd, synth, synth.planar_star(n=N))
    elif leg == "reference_example":
        out = leg_reference_example(nd, synth)
    elif leg == "accuracy":
        out = leg_accuracy_planar(nd, synth, synth.planar_star(n=N))
    elif leg == "capacity":
        serve = importlib.import_module("nadia-coecxs_amd.serve")
        reqs = serve.poisson_stream(count=args.serve_requests, rate=args.serve_rate)
        sizes = sorted({r.n for r in reqs})
        be = serve.DeviceBackend(nd, serve.input_pool(synth, sizes, per_size=2))
        budget = args.serve_budget_gb * 1e9
        serve.Server(be, hbm_budget=budget).run(
            [serve.Request(-1 - i, 0.0, n, serve.HIO, 0.9, 20, False, 0) for i, n in enumerate(sizes)])
        out = leg_capacity(nd, serve, reqs, be, budget)
    else:
        raise SystemExit("unknown leg %r" % leg)
    print(json.dumps({leg: out}))


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
    ap.add_argument("--no-accuracy", action="store_true", help="skip the FP32-vs-FP64 schedule runs")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-job batch")
    ap.add_argument("--batched-iters", type=int, default=500)
    ap.add_argument("--graph-steps", type=int, default=50)
    ap.add_argument("--no-fresnel", action="store_true", help="skip the 1024^2 Fresnel workload")
    ap.add_argument("--no-served", action="store_true", help="skip the request stream")
    ap.add_argument("--no-phase-diverse", action="store_true", help="skip the PhaseDiverseCDI leg")
    ap.add_argument("--no-next-rows", action="store_true",
                    help="skip the constraint / n-best / autocorrelation leg")
    ap.add_argument("--no-reference-example", action="store_true",
                    help="skip the reference's PlanarCDI example (measured 1024^2 pattern)")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-rate", type=float, default=40.0)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--profile", default=None, metavar="LEG",
                    choices=["batched", "fresnel", "phase_diverse", "next_rows", "capacity", "accuracy",
                             "reference_example"],
                    help="run one leg alone (for profiler runs) and print its JSON")
    ap.add_argument("--copy-bandwidth", action="store_true",
                    help="device copy rate (read + write) at 96 MB .. 4 GB, then exit")
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
    if args.copy_bandwidth:
        # the device's own copy rate (torch's copy kernel on its stream, events
        # on that stream): source + destination inside the 256 MB Infinity
        # Cache, then streaming from HBM
        import torch
        out = {}
        for mb in (96, 256, 1024, 4096):
            n = mb * (1 << 20) // 4
            a = torch.ones(n, dtype=torch.float32, device="cuda")
            b = torch.empty_like(a)
            for _ in range(3):
                b.copy_(a)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(20):
                b.copy_(a)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / 20
            out["copy_%dMB_TBs" % mb] = round(2 * n * 4 / ms / 1e9, 2)
            del a, b
        print(json.dumps({"copy_bandwidth": out}))
        return
    if args.copy_calibration:
        a = nd.Complex2D(N, N, double=False)
        b = nd.Complex2D(N, N, double=False)
        for _ in range(args.copy_calibration):
            b.copy(a)
        nd.synchronize()
        print(json.dumps({"copy_calibration": args.copy_calibration, "bytes_read_per_copy": N * N * 8,
                          "bytes_written_per_copy": N * N * 8}))
        return
    if args.profile:
        profile_leg(nd, synth, args)
        return
    t_gen = time.perf_counter()
    data = synth.planar_star(n=N)
    t_gen = time.perf_counter() - t_gen
    print("bench: inputs generated in %.1f s" % t_gen, file=sys.stderr, flush=True)

    # ---- FP32 fused, batched (the headline value) ----------------------
    # one job alone, then two independent 2048^2 jobs sharing every launch;
    # value = the two jobs' aggregate job-iterations per second
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
    # the metric is the batched FP32 rate: two 2048^2 jobs sharing every
    # launch (the single-job rate is reported beside it)
    batch_jobs = 2
    value = dual
    # eight jobs: 8 x 96 MB of hot state, well beyond the 256 MB Infinity
    # Cache that holds one or two jobs -- the rate when every pass streams
    # from HBM
    jobs8 = [setup(nd, data, False, nd.PATH_FUSED) for _ in range(8)]
    batch = nd.Batch([q for _, q in jobs8])
    batch.iterate_async(max(2, args.warmup // 4))
    nd.synchronize()
    n8 = max(10, args.steps // 4)
    t0 = time.perf_counter()
    batch.iterate_async(n8)
    nd.synchronize()
    batch8 = 8 * n8 / (time.perf_counter() - t0)
    print("bench: fp32 fused, batch of 8 jobs: %.1f job-iter/s" % batch8, file=sys.stderr, flush=True)
    del batch, jobs8
    # per-pass device time with HIP events on the library stream (one job)
    p.set_timing(True)
    p.iterate_async(min(args.steps, 200))
    nd.synchronize()
    ms, cnt = p.pass_times()
    p.set_timing(False)
    ev = {"col": ms[0] / max(cnt[0], 1), "row": ms[1] / max(cnt[1], 1)}
    # Per-launch events add ~1 us each; the pass durations are the iteration
    # time of the timed run (no per-launch events) split in the ratio of the
    # per-launch event times.
    t_iter_ms = 1000.0 / single
    avg = {k: t_iter_ms * v / (ev["col"] + ev["row"]) for k, v in ev.items()}
    b32 = pass_bytes(False)
    # the longer pass; when the two are within 3% of each other (the usual
    # case at 2048^2), the column pass, which is further from its roofline
    dom = "col" if avg["col"] >= 0.97 * avg["row"] else "row"
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
        "col_pass_frac": round(b32["col"] * N * N / (avg["col"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "row_pass_frac": round(b32["row"] * N * N / (avg["row"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "col_pass_event_ms": round(ev["col"], 5), "row_pass_event_ms": round(ev["row"], 5),
        "tile_variants": list(p.get_tile_variants()),
        # the scope's count for an HIO iteration (7 complex + 2 real per pixel)
        # and the bytes the two fused passes move (6 complex + 2 real)
        "iteration_effective_GBs": round(b32["iter"] * N * N * single / 1e9, 1),
        "iteration_moved_GBs": round((b32["col"] + b32["row"]) * N * N * single / 1e9, 1),
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

    extra = {}
    if not args.no_accuracy:
        extra["accuracy_2048"] = leg_accuracy_planar(nd, synth, data)
        print("bench: accuracy %s" % json.dumps(extra["accuracy_2048"]), file=sys.stderr, flush=True)
    if not args.no_batched:
        extra["batched64"] = leg_batched(nd, synth, args.batched_iters, args.graph_steps)
        print("bench: batched64 %s" % json.dumps(extra["batched64"]), file=sys.stderr, flush=True)
    if not args.no_fresnel:
        extra["fresnel1024"] = leg_fresnel(nd, synth)
        print("bench: fresnel %s" % json.dumps(extra["fresnel1024"]), file=sys.stderr, flush=True)
    if not args.no_reference_example:
        extra["reference_example_1024"] = leg_reference_example(nd, synth)
        print("bench: reference example %s" % json.dumps(extra["reference_example_1024"]), file=sys.stderr,
              flush=True)
    if not args.no_phase_diverse:
        extra["phase_diverse"] = leg_phase_diverse(nd, synth)
        print("bench: phase diverse %s" % json.dumps(extra["phase_diverse"]), file=sys.stderr, flush=True)
    if not args.no_next_rows:
        extra["next_rows"] = leg_next_rows(nd, synth, data)
        print("bench: next rows %s" % json.dumps(extra["next_rows"]), file=sys.stderr, flush=True)
    if not args.no_served:
        extra["served"] = leg_served(nd, synth, args.serve_requests, args.serve_rate, args.serve_budget_gb * 1e9)
        print("bench: served %s" % json.dumps(extra["served"]), file=sys.stderr, flush=True)

    # ---- CPU baseline: the oracle on the host, bounded sample -----------
    # the oracle's timing builds, compiled with the reference's flags
    # (-O3 -fPIC -ffast-math): FP32 with its 2-D transforms over T threads
    # (NADIA's default libNADIA threads only its FFTW calls) and on one
    # thread; FP64 on one thread (the threaded FP64 build of the reference
    # does not compile)
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle  # CPU restatement of the reference (checker / baseline only)

        nproc = os.cpu_count() or 1
        try:
            affinity = len(os.sched_getaffinity(0))
        except AttributeError:
            affinity = nproc
        threads = int(os.environ.get("OMP_NUM_THREADS") or affinity)
        try:
            gomp = ctypes.CDLL("libgomp.so.1")
        except OSError:
            gomp = None

        def time_oracle(double, nthreads, iters):
            if gomp is not None:
                gomp.omp_set_num_threads(nthreads)
            cdt = np.complex128 if double else np.complex64
            xo = data["x0"].astype(cdt)
            o = oracle.OracleCDI(xo, double=double, timing=True)
            o.set_support(data["support"])
            o.set_intensity(data["intensity"])
            o.set_beam_stop(data["beam_stop"])
            o.set_algorithm(oracle.HIO)
            o.iterate()  # warm-up (temporaries, first touch)
            t0 = time.perf_counter()
            o.iterate_n(iters)
            r = iters / (time.perf_counter() - t0)
            print("bench: cpu oracle %s %d thread(s) %.3f iter/s" % ("fp64" if double else "fp32", nthreads, r),
                  file=sys.stderr, flush=True)
            return r

        n1 = max(2, args.cpu_sample // 2)
        mt = time_oracle(False, threads, args.cpu_sample * 2)
        st = time_oracle(False, 1, n1)
        d1 = time_oracle(True, 1, n1)
        cpu = {"value": round(mt, 4), "unit": "iter/s", "cores": threads, "kind": "port",
               "sample": "%d HIO iterations at 2048^2 after 1 warm-up, FP32 arrays with FP64 error sums (the "
                         "reference's default precision), oracle built with the reference's -O3 -ffast-math, its "
                         "2-D transforms over %d OpenMP threads (the oracle's own FFT; no FFTW on the host)"
                         % (args.cpu_sample * 2, threads),
               "fp32_1thread_iter_s": round(st, 4), "fp64_1thread_iter_s": round(d1, 4),
               "nproc": nproc, "affinity_cpus": affinity, "cpu_model": cpu_model()}

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "iter/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1000.0 / value, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "PlanarCDI HIO beta=0.9, 2048x2048 star-shaped complex object, "
                               "Poisson peak 1e8, 16x16 beam stop, 900x900 support, FP32 bulk + FP64 error sums",
                   "nx": N, "ny": N, "algorithm": "HIO", "beta": 0.9, "batch_jobs": batch_jobs},
        "fp32_single_job_iter_s": round(single, 2), "fp32_batch2_job_iter_s": round(dual, 2),
        "fp32_batch8_job_iter_s": round(batch8, 2),
        "fp32_batch8_moved_GBs": round(batch8 * (pass_bytes(False)["col"] + pass_bytes(False)["row"]) * N * N / 1e9, 1),
        "fp64_plain_iter_s": round(plain64, 2), "fp64_fused_iter_s": round(fused64, 2),
        "fp32_vs_fp64_plain": round(value / plain64, 2),
        "final_error_fp32": err32,
        "roofline": roofline, "cpu_baseline": cpu,
    }
    out.update(extra)
    if "served" in extra:
        out["served_p50_ms"] = extra["served"]["p50_ms"]
        out["served_p99_ms"] = extra["served"]["p99_ms"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
