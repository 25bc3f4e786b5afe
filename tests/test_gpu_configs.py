"""GPU parity at the sizes of the benchmark workloads:

* FresnelCDI 1024^2 with the zone-plate geometry of the Fresnel workload
  (lambda 1.55e-10 m, focal-sample 1.6 mm, focal-detector 0.909 m, 13.5 um
  pixels): one FP64 ER iteration <= 1e-12 relative L2 against the oracle,
  and one FresnelCDI_WF white-field iteration likewise; FP32 over 5
  iterations within 1e-4 of the FP64 oracle;
* the 64-job RAAR batch (sizes 256..1024) in graph mode: every job
  bit-identical to the same job iterated alone, and one FP64 job of each
  size <= 1e-12 against the oracle for one iteration;
* a served stream over 256^2..2048^2 with ER / HIO / RAAR, shrinkwrap on and
  off and requests with and without a beam stop, admitted a few at a time by
  the memory budget (idle reconstructions are reused): each result equals the
  same job run alone;
* a slice of the 2048^2 HIO + shrinkwrap schedule in FP64: 20 HIO
  iterations, one shrinkwrap, 5 ER; supports bit-equal to the oracle's,
  errors <= 1e-9 over the 20 HIO iterations.
"""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

serve = importlib.import_module("nadia-coecxs_amd.serve")


def rel(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# ------------------------------------------------------ Fresnel 1024^2 ----

@pytest.fixture(scope="module")
def fresnel1024(synth):
    return synth.fresnel_config(n=1024)


def fresnel_pair(nd, oracle, d, double):
    n = 1024
    cdt = np.complex128 if double else np.complex64
    est = nd.Complex2D(n, n, double=double)
    g = nd.FresnelCDI(est, d["white_field"].astype(cdt), d["wavelength"], d["zfd"], d["zfs"], d["pixel"], 1.0)
    xo = np.zeros((n, n), np.complex128)
    o = oracle.OracleCDI(xo, oracle.FRESNEL)
    o.fresnel_init(d["white_field"], d["wavelength"], d["zfd"], d["zfs"], d["pixel"], 1.0)
    for obj in (g, o):
        obj.set_support(d["support"])
        obj.set_intensity(d["intensity"])
        obj.initialise_estimate(0)
    return est, g, xo, o


def wf_pair(nd, oracle, d, double):
    n = 1024
    est = nd.Complex2D(n, n, double=double)
    g = nd.FresnelCDI_WF(est, d["wavelength"], d["z12"], d["z23"], d["pixel"])
    g.set_support(d["pupil_diameter"])
    xo = np.zeros((n, n), np.complex128)
    o = oracle.OracleCDI(xo, oracle.FRESNEL_WF)
    o.wf_init(d["wavelength"], d["z12"], d["z23"], d["pixel"])
    o.wf_set_support(d["pupil_diameter"])
    for obj in (g, o):
        obj.set_intensity(d["white_field_intensity"])
        obj.initialise_estimate(0)
    return est, g, xo, o


@pytest.mark.parametrize("stage", ["fresnel_er", "white_field"])
def test_fresnel1024_fp64_one_iteration(nd, oracle, fresnel1024, stage):
    pair = fresnel_pair if stage == "fresnel_er" else wf_pair
    est, g, xo, o = pair(nd, oracle, fresnel1024, True)
    assert rel(est.get(), xo) <= 1e-14
    for k in range(2):
        e = g.iterate(1)[0]
        o.iterate()
        assert rel(est.get(), xo) <= (1e-12 if k == 0 else 1e-11), k
        assert abs(e - o.get_error()) <= 1e-12 * o.get_error(), k


@pytest.mark.parametrize("stage", ["fresnel_er", "white_field"])
def test_fresnel1024_fp32_five_iterations(nd, oracle, fresnel1024, stage):
    pair = fresnel_pair if stage == "fresnel_er" else wf_pair
    est, g, _, _ = pair(nd, oracle, fresnel1024, False)
    _, _, xo, o = pair(nd, oracle, fresnel1024, True)
    e = g.iterate(5)
    eo = o.iterate_n(5)
    assert rel(est.get(), xo) <= 1e-4
    assert (np.abs(e - eo) / eo).max() <= 1e-4


# ------------------------------------------------------ 64-job batch ----

@pytest.fixture(scope="module")
def jobs64(synth):
    return synth.batched_jobs()


def raar_job(nd, d, double=False):
    n = d["n"]
    cdt = np.complex128 if double else np.complex64
    est = nd.Complex2D(n, n, double=double, data=d["x0"].astype(cdt))
    g = nd.PlanarCDI(est)
    g.set_support(d["support"])
    g.set_intensity(d["intensity"])
    g.set_relaxation_parameter(0.85)
    g.set_algorithm(nd.RAAR)
    return est, g


def test_batch64_graph_bit_identical_to_alone(nd, jobs64):
    assert len(jobs64) == 64 and {d["n"] for d in jobs64} == {256, 384, 512, 768, 1024}
    js = [raar_job(nd, d) for d in jobs64]
    b = nd.Batch([g for _, g in js])
    b.set_graph(5)
    e = b.iterate(10)
    assert b.graph_replays() == 2
    for j, d in enumerate(jobs64):
        est, g = raar_job(nd, d)
        ea = g.iterate(10)
        assert np.array_equal(js[j][0].get(), est.get()), j
        assert np.array_equal(e[:, j], ea), j


def test_batch64_fp64_each_size_matches_oracle(nd, oracle, jobs64):
    seen = set()
    for d in jobs64:
        if d["n"] in seen:
            continue
        seen.add(d["n"])
        est, g = raar_job(nd, d, double=True)
        xo = d["x0"].astype(np.complex128).copy()
        o = oracle.OracleCDI(xo)
        o.set_support(d["support"])
        o.set_intensity(d["intensity"])
        o.set_relaxation_parameter(0.85)
        o.set_algorithm(oracle.RAAR)
        e = g.iterate(1)[0]
        o.iterate()
        assert rel(est.get(), xo) <= 1e-12, d["n"]
        assert abs(e - o.get_error()) <= 1e-12 * o.get_error(), d["n"]
    assert len(seen) == 5


# ------------------------------------------------------ served stream ----

def test_served_mix_256_to_2048_matches_alone(nd, synth):
    sizes = (256, 512, 1024, 2048)
    pool = {n: [synth.planar_star(n=n, seed=60 + k) for k in range(1 if n == 2048 else 2)] for n in sizes}

    def inputs(r):
        d = dict(pool[r.n][r.seed % len(pool[r.n])])
        if r.rid % 5 == 0:
            del d["beam_stop"]
        return d

    algs = ((serve.ER, 0.9), (serve.HIO, 0.9), (serve.RAAR, 0.85))
    reqs = [serve.Request(i, 0.0005 * i, sizes[i % 4], algs[i % 3][0], algs[i % 3][1], 60 + 17 * i,
                          (i // 2) % 2 == 0, i) for i in range(14)]
    assert {r.algorithm for r in reqs} == {serve.ER, serve.HIO, serve.RAAR}
    assert any(r.shrinkwrap for r in reqs) and not all(r.shrinkwrap for r in reqs)
    be = serve.DeviceBackend(nd, inputs)
    budget = nd.bytes_estimate(2048, 2048) + 2 * nd.bytes_estimate(1024, 1024)
    srv = serve.Server(be, hbm_budget=budget, tick=40, keep_results=True)
    done = srv.run(reqs)
    assert len(done) == 14 and srv.peak_bytes <= budget and srv.peak_active >= 2
    for c in done:
        r = reqs[c.rid]
        d = inputs(r)
        est = nd.Complex2D(r.n, r.n, double=False, data=d["x0"].astype(np.complex64))
        g = nd.PlanarCDI(est)
        g.set_support(d["support"])
        g.set_intensity(d["intensity"])
        if "beam_stop" in d:
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


# --------------------------------------------- 2048^2 HIO + shrinkwrap ----

def test_hio2048_shrinkwrap_slice_fp64(nd, oracle, synth):
    n = 2048
    d = synth.planar_star(n=n)
    est = nd.Complex2D(n, n, double=True, data=d["x0"])
    g = nd.PlanarCDI(est)
    xo = d["x0"].astype(np.complex128).copy()
    o = oracle.OracleCDI(xo)
    for obj in (g, o):
        obj.set_support(d["support"])
        obj.set_intensity(d["intensity"])
        obj.set_beam_stop(d["beam_stop"])
        obj.set_algorithm(nd.HIO)
    e = g.iterate(20)
    eo = o.iterate_n(20)
    assert (np.abs(e - eo) / eo).max() <= 1e-9
    assert rel(est.get(), xo) <= 1e-9
    g.apply_shrinkwrap(3.0, 0.1)
    o.apply_shrinkwrap(3.0, 0.1)
    assert np.array_equal(g.get_support(), o.get_support())
    for obj in (g, o):
        obj.set_algorithm(nd.ER)
    e = g.iterate(5)
    eo = o.iterate_n(5)
    assert (np.abs(e - eo) / eo).max() <= 1e-8
    assert rel(est.get(), xo) <= 1e-8
