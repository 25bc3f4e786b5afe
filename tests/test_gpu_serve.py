"""Per-job iteration counts, graph-captured batch sweeps and the request
loop on the device: every job's result is bit-identical to the same job
iterated alone (tolerance 0: batched launches compute each job with the same
kernels, tile variants and per-job reductions)."""
import importlib

import numpy as np
import pytest

from test_gpu_batch import planar_job

pytestmark = pytest.mark.gpu

serve = importlib.import_module("nadia-coecxs_amd.serve")


def test_iterate_each_matches_alone(nd, synth):
    spec = [(256, 21, nd.HIO, False, 9), (256, 22, nd.RAAR, False, 4), (256, 23, nd.ER, False, 0),
            (512, 24, nd.HIO, False, 7), (384, 25, nd.DM, True, 3), (256, 26, nd.HIO, True, 5)]
    alone = []
    for n, seed, alg, double, cnt in spec:
        est, g = planar_job(nd, synth, n, seed, alg, double)
        if cnt:
            g.iterate(cnt)
        alone.append((est.get(), g.get_error() if cnt else None))
    jobs = [planar_job(nd, synth, n, seed, alg, double) for n, seed, alg, double, _ in spec]
    b = nd.Batch([g for _, g in jobs])
    counts = [c for *_, c in spec]
    # two calls: the second continues where the first stopped
    first = [min(c, 3) for c in counts]
    b.iterate_each(first)
    b.iterate_each([c - f for c, f in zip(counts, first)])
    for j, (est, g) in enumerate(jobs):
        x0, e0 = alone[j]
        assert np.array_equal(est.get(), x0), j
        if e0 is not None:
            assert g.get_error() == e0, j


@pytest.mark.parametrize("double", [False, True], ids=["fp32", "fp64"])
def test_graph_sweeps_match_eager(nd, synth, double):
    """Graph replays run the launch groups as parallel branches; every job
    ends bit-identical to eager serial launches.  The 256^2 jobs form two
    groups of one size (one and two work fields), the case that once showed
    FP64 differences from a kernel store hazard under concurrency."""
    spec = [(256, 31, nd.RAAR), (384, 32, nd.RAAR), (512, 33, nd.HIO), (256, 34, nd.ER), (768, 35, nd.RAAR),
            (256, 36, nd.DM)]
    eager = [planar_job(nd, synth, n, s, a, double=double) for n, s, a in spec]
    be = nd.Batch([g for _, g in eager])
    e_eager = np.vstack([be.iterate(25), be.iterate(25), be.iterate(7)])
    graph = [planar_job(nd, synth, n, s, a, double=double) for n, s, a in spec]
    bg = nd.Batch([g for _, g in graph])
    bg.set_graph(10)
    # eager host work between graph replays (downloads, another batch)
    e1 = bg.iterate(25)      # 2 replays + 5 eager iterations
    _ = [est.get() for est, _ in graph]
    e2 = bg.iterate(25)      # 2 more replays of the same graph
    e3 = bg.iterate(7)       # shorter than the graph: eager
    assert bg.graph_replays() == 4
    assert np.array_equal(np.vstack([e1, e2, e3]), e_eager)
    for (xe, ge), (xg, gg) in zip(eager, graph):
        assert np.array_equal(xe.get(), xg.get())
        assert ge.get_error() == gg.get_error()


def test_device_bytes(nd, synth):
    est, g = planar_job(nd, synth, 256, 41, nd.HIO, double=False, beam=True)
    g.iterate(2)
    # the fused path allocates no unfused temporaries
    base = g.device_bytes()
    assert base + 256 * 256 * 8 <= nd.bytes_estimate(256, 256, double=False)
    g.set_path(nd.PATH_PLAIN)
    g.iterate(1)
    assert g.device_bytes() > base


def test_served_results_match_alone(nd, synth):
    """Requests arriving together, admitted a few at a time by the memory
    budget; each result equals the same job run alone."""
    sizes = (256, 512)
    pool = {n: [synth.planar_star(n=n, seed=50 + k) for k in range(2)] for n in sizes}
    inputs = lambda r: pool[r.n][r.seed % 2]
    algs = ((serve.ER, 0.9), (serve.HIO, 0.9), (serve.RAAR, 0.85))
    reqs = [serve.Request(i, 0.0005 * i, sizes[i % 2], algs[i % 3][0], algs[i % 3][1], 60 + 23 * i, i % 2 == 0, i)
            for i in range(8)]
    be = serve.DeviceBackend(nd, inputs)
    budget = 3 * nd.bytes_estimate(512, 512)
    srv = serve.Server(be, hbm_budget=budget, tick=40, keep_results=True)
    done = srv.run(reqs)
    assert len(done) == 8 and srv.peak_bytes <= budget
    assert srv.peak_active >= 2
    for c in done:
        r = reqs[c.rid]
        d = inputs(r)
        est = nd.Complex2D(r.n, r.n, double=False, data=d["x0"].astype(np.complex64))
        g = nd.PlanarCDI(est)
        g.set_support(d["support"])
        g.set_intensity(d["intensity"])
        g.set_beam_stop(d["beam_stop"])
        g.set_relaxation_parameter(r.beta)
        g.set_algorithm(r.algorithm)
        ncalls = (r.iterations - 1) // 50 if r.shrinkwrap else 0
        it = 0
        for k in range(ncalls):
            g.iterate_async(50 * (k + 1) - it)
            it = 50 * (k + 1)
            g.apply_shrinkwrap(serve.shrinkwrap_sigma(k, ncalls), 0.1)
        g.iterate_async(r.iterations - it)
        assert np.array_equal(est.get(), c.result), c.rid
        assert g.get_error() == c.error, c.rid


def test_event_and_async_error(nd, synth):
    """get_error copied out asynchronously into page-locked memory equals
    get_error() once an event recorded after it has completed."""
    est, g = planar_job(nd, synth, 256, 45, nd.HIO, double=False)
    buf = nd.PinnedDoubles(4)
    ev = nd.Event()
    assert ev.query()  # never recorded: nothing to wait for
    g.iterate_async(30)
    g.get_error_async(buf, 2)
    ev.record()
    ev.synchronize()
    assert ev.query()
    assert buf.a[2] == g.get_error() and 0 < buf.a[2] < 1
