"""Batched iteration (nadia_batch): every job's estimate and errors are
bit-identical to iterating the same job alone, whatever else is in the
batch (uneven sizes, kinds, algorithms, precisions)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def planar_job(nd, synth, n, seed, alg, double=True, beta=0.9, beam=False, n_best=0):
    d = synth.planar_block(n=n, block=3 * n // 8, support=7 * n // 16, seed_amp=seed, seed_poisson=seed,
                           seed_phase=seed + 1)
    cdt = np.complex128 if double else np.complex64
    est = nd.Complex2D(n, n, double=double, data=d["x0"].astype(cdt))
    g = nd.PlanarCDI(est, n_best=n_best)
    g.set_support(d["support"])
    g.set_intensity(d["intensity"])
    if beam:
        b = np.ones((n, n))
        b[n // 2 - 2:n // 2 + 2, n // 2 - 2:n // 2 + 2] = 0
        g.set_beam_stop(b)
    g.set_relaxation_parameter(beta)
    g.set_algorithm(alg)
    return est, g


def specs(nd):
    return [(256, 1, nd.HIO, True), (256, 2, nd.RAAR, True), (384, 3, nd.ER, True), (256, 4, nd.DM, True),
            (512, 5, nd.HIO, False), (256, 6, nd.HIO, False), (768, 7, nd.RAAR, False)]


def test_batch_matches_alone(nd, synth):
    alone = []
    for n, seed, alg, double in specs(nd):
        est, g = planar_job(nd, synth, n, seed, alg, double, beam=seed % 2 == 0)
        e = g.iterate(7)
        alone.append((est.get(), e))
    jobs = [planar_job(nd, synth, n, seed, alg, double, beam=seed % 2 == 0) for n, seed, alg, double in specs(nd)]
    b = nd.Batch([g for _, g in jobs])
    e = np.vstack([b.iterate(3), b.iterate(4)])
    assert e.shape == (7, len(jobs))
    for j, (est, g) in enumerate(jobs):
        x0, e0 = alone[j]
        assert np.array_equal(est.get(), x0), j
        assert np.array_equal(e[:, j], e0), j
        assert g.get_error() == e0[-1]


def test_batch_membership_changes(nd, synth):
    """Jobs join and leave between iterations (continuous batching): a job's
    trajectory does not depend on when others come and go."""
    est_a, a = planar_job(nd, synth, 256, 11, nd.HIO)
    ref_est, ref = planar_job(nd, synth, 256, 11, nd.HIO)
    ref_e = ref.iterate(6)
    b = nd.Batch([a])
    e = list(b.iterate(2)[:, 0])
    est_b, bjob = planar_job(nd, synth, 256, 12, nd.RAAR)
    b.add(bjob)
    e += list(b.iterate(2)[:, 0])
    b.remove(bjob)
    e += list(b.iterate(2)[:, 0])
    assert np.array_equal(np.array(e), ref_e)
    assert np.array_equal(est_a.get(), ref_est.get())


def test_batch_fresnel_and_wf(nd, synth):
    d = synth.fresnel_config(n=128, pupil_diameter=160e-6 / 8, disk_radius=4.5, ndisks=3)
    def fres():
        est = nd.Complex2D(128, 128)
        g = nd.FresnelCDI(est, d["white_field"], d["wavelength"], d["zfd"], d["zfs"], d["pixel"], 1.0)
        g.set_support(d["support"])
        g.set_intensity(d["intensity"])
        g.initialise_estimate(0)
        return est, g
    def wf():
        est = nd.Complex2D(128, 128)
        g = nd.FresnelCDI_WF(est, d["wavelength"], d["z12"], d["z23"], d["pixel"])
        g.set_support(d["pupil_diameter"])
        g.set_intensity(d["white_field_intensity"])
        g.initialise_estimate(0)
        return est, g
    (e1, g1), (e2, g2) = fres(), wf()
    r1, r2 = g1.iterate(5), g2.iterate(5)
    (f1, h1), (f2, h2) = fres(), wf()
    b = nd.Batch([h1, h2])
    e = b.iterate(5)
    assert np.array_equal(e[:, 0], r1) and np.array_equal(e[:, 1], r2)
    assert np.array_equal(f1.get(), e1.get()) and np.array_equal(f2.get(), e2.get())


@pytest.mark.parametrize("graph", [0, 4], ids=["eager", "graph"])
def test_batch_n_best_matches_alone(nd, synth, graph):
    """Best estimates kept on the device inside batched launches (eager and
    graph replays) equal those of the same jobs iterated alone."""
    spec = [(256, 61, nd.HIO), (256, 62, nd.ER), (128, 63, nd.RAAR)]
    alone = []
    for n, seed, alg in spec:
        est, g = planar_job(nd, synth, n, seed, alg, n_best=3)
        g.iterate(9)
        alone.append([g.get_best_result(k) for k in range(3)])
    jobs = [planar_job(nd, synth, n, seed, alg, n_best=3) for n, seed, alg in spec]
    b = nd.Batch([g for _, g in jobs])
    if graph:
        b.set_graph(graph)
    b.iterate(9)
    for (est, g), want in zip(jobs, alone):
        for k in range(3):
            x, e = g.get_best_result(k)
            assert np.array_equal(x, want[k][0]) and e == want[k][1], k


@pytest.mark.parametrize("double", [True, False], ids=["fp64", "fp32"])
def test_shared_launch_variant_timing_keeps_results(nd, synth, double):
    """A batch that is one launch group times the column tile variants on
    the first 18 iterations of a call (blocks of three iterations per
    variant) and uses the fastest from a later call on: every job's estimate
    and errors stay bit-identical to the job run alone, through the timed
    call, the call that settles the choice and the calls after it."""
    counts = (20, 20, 7)
    alone = []
    for seed in (61, 62, 63):
        est, g = planar_job(nd, synth, 512, seed, nd.HIO, double)
        e = np.concatenate([g.iterate(c) for c in counts])
        alone.append((est.get(), e))
    jobs = [planar_job(nd, synth, 512, seed, nd.HIO, double) for seed in (61, 62, 63)]
    b = nd.Batch([g for _, g in jobs])
    e = np.vstack([b.iterate(c) for c in counts])
    for j, (est, g) in enumerate(jobs):
        x0, e0 = alone[j]
        assert np.array_equal(est.get(), x0), j
        assert np.array_equal(e[:, j], e0), j


def test_rectangular_jobs_in_a_batch_match_alone(nd, synth):
    """Jobs of transposed rectangular shapes (their own launch groups) and a
    square one, three algorithms, FP64 and FP32: each bit-identical to the
    same job run alone."""
    def job(shape, seed, alg, double):
        nx, ny = shape
        rng = np.random.default_rng(seed)
        obj = np.zeros(shape, complex)
        obj[nx // 4:3 * nx // 4, ny // 4:3 * ny // 4] = rng.uniform(0.5, 1, (nx // 2, ny // 2))
        I = np.abs(np.fft.fft2(obj)) ** 2 + 1e-3
        sup = np.zeros(shape)
        sup[nx // 8:7 * nx // 8, ny // 8:7 * ny // 8] = 1.0
        cdt = np.complex128 if double else np.complex64
        est = nd.Complex2D(nx, ny, double=double, data=(sup * np.exp(2j * np.pi * rng.random(shape))).astype(cdt))
        g = nd.PlanarCDI(est)
        g.set_support(sup)
        g.set_intensity(I)
        g.set_algorithm(alg)
        return est, g

    spec = [((256, 512), 71, nd.HIO, True), ((512, 256), 72, nd.RAAR, True), ((256, 512), 73, nd.ER, True),
            ((256, 256), 74, nd.HIO, False), ((128, 768), 75, nd.HIO, False), ((768, 128), 76, nd.ER, False)]
    alone = []
    for shape, seed, alg, double in spec:
        est, g = job(shape, seed, alg, double)
        e = g.iterate(9)
        alone.append((est.get(), e))
    jobs = [job(*s) for s in spec]
    b = nd.Batch([g for _, g in jobs])
    e = np.vstack([b.iterate(4), b.iterate(5)])
    for j, (est, g) in enumerate(jobs):
        x0, e0 = alone[j]
        assert np.array_equal(est.get(), x0), j
        assert np.array_equal(e[:, j], e0), j
