"""GPU parity of PhaseDiverseCDI position refinement (nadia_pd_adjust_positions,
nadia_align) with the oracle (NADIA/src/PhaseDiverseCDI.c++:337-719,
NADIA/src/utils.c:1151-1435): offsets found by align, positions after
adjust_positions in both search types and directions, the frames' estimates
afterwards, and the untouched global object."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

synth = importlib.import_module("nadia-coecxs_amd.synth")


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("double", [True, False], ids=["fp64", "fp32"])
@pytest.mark.parametrize("seed,window,weights", [(1, (0, 0, 0, 0), False), (2, (-5, 5, -4, 6), True),
                                                 (3, (-3, 9, -9, 2), True)])
def test_align_matches_oracle(nd, oracle, double, seed, window, weights):
    rng = np.random.default_rng(seed)
    a = rng.uniform(0, 1, (16, 24))
    b = np.roll(a, (2, -3), axis=(0, 1)) + 0.05 * rng.uniform(0, 1, a.shape)
    wa = (rng.uniform(0, 1, a.shape) > 0.2).astype(float) if weights else None
    wb = (rng.uniform(0, 1, b.shape) > 0.3).astype(float) if weights else None
    got = nd.align(a, b, *window, w1=wa, w2=wb, double=double)
    want = oracle.oracle_align(a, b, *window, wa=wa, wb=wb)
    assert got == want


@pytest.mark.parametrize("shift", [(3, -5), (-7, 2)])
def test_align_recovers_a_shift(nd, shift):
    g = synth.disk_field(160, 60, np.random.default_rng(4))
    t = shift
    a = g[40:104, 40:104]
    b = g[40 - t[0]:104 - t[0], 40 - t[1]:104 - t[1]]
    assert nd.align(a, b, -10, 10, -10, 10) == (True, -t[0], -t[1])


def build(nd, oracle, d, starts, double=True, parallel=False, granularity=1):
    cdt = np.complex128 if double else np.complex64
    ests, gs, xs, os_ = [], [], [], []
    for f, inten, sup in zip(d["frames"], d["intensities"], d["supports"]):
        est = nd.Complex2D(f.shape[0], f.shape[1], double=double, data=f.astype(cdt))
        g = nd.PlanarCDI(est)
        x = f.astype(np.complex128).copy()
        o = oracle.OracleCDI(x, oracle.PLANAR)
        for obj in (g, o):
            obj.set_support(sup)
            obj.set_intensity(inten)
        g.set_algorithm(nd.ER)
        o.set_algorithm(oracle.ER)
        ests.append(est)
        gs.append(g)
        xs.append(x)
        os_.append(o)
    pg = nd.PhaseDiverseCDI(beta=1.0, parallel=parallel, granularity=granularity)
    po = oracle.OraclePhaseDiverse(beta=1.0, parallel=parallel, granularity=granularity)
    for g, o, (x0, y0) in zip(gs, os_, starts):
        pg.add_new_position(g, x0, y0)
        po.add_new_position(o, x0, y0)
    pg.initialise_estimate()
    po.initialise_estimate()
    return pg, po, ests, xs, (gs, os_)


@pytest.mark.parametrize("forward", [True, False], ids=["forward", "backward"])
@pytest.mark.parametrize("kind", [0, 1], ids=["cross_correlation", "minimum_error"])
@pytest.mark.parametrize("parallel", [False, True], ids=["series", "parallel"])
def test_adjust_positions_matches_oracle(nd, oracle, kind, forward, parallel):
    d = synth.ptycho_frames()
    truth = d["positions"]
    ref = 0 if forward else len(truth) - 1
    starts = [(x + (0 if k == ref else 2.0 * (-1) ** k), y + (0 if k == ref else -3.0 + k))
              for k, (x, y) in enumerate(truth)]
    pg, po, ests, xs, keep = build(nd, oracle, d, starts, parallel=parallel)
    before = pg.get_transmission()
    gg = pg.adjust_positions(kind, forward, step_size=4.0)
    go = po.adjust_positions(kind, forward, step_size=4.0)
    assert np.array_equal(gg, go)
    for k in range(len(truth)):
        assert pg.get_position(k) == po.get_position(k)
        assert pg.get_position(k) == (truth[k][0] - truth[ref][0] + starts[ref][0],
                                      truth[k][1] - truth[ref][1] + starts[ref][1])
        assert rel(ests[k].get(), xs[k]) <= 1e-10
    assert np.array_equal(pg.get_transmission(), before)
    assert rel(before, po.get_transmission()) <= 1e-12
    # the refined positions carry on into the next big iteration: update
    # from and add to the object at the new positions (no frame iteration --
    # after a frame has moved its merged estimate can have a Fourier
    # amplitude of exactly zero where sqrt(I) > 0, where the modulus
    # projection's phase is set by rounding on either implementation)
    pg.set_iterations_per_cycle(0)
    po.set_iterations_per_cycle(0)
    pg.iterate()
    po.iterate()
    assert rel(pg.get_transmission(), po.get_transmission()) <= 1e-12
    for k in range(len(truth)):
        assert rel(ests[k].get(), xs[k]) <= 1e-12


def test_adjust_positions_fp32_and_granularity_2(nd, oracle):
    d = synth.ptycho_frames()
    truth = d["positions"]
    starts = [(x + (0 if k == 0 else 1.0), y + (0 if k == 0 else -2.0)) for k, (x, y) in enumerate(truth)]
    # FP32 device frames against the FP64 oracle: the same positions
    pg, po, ests, xs, keep = build(nd, oracle, d, starts, double=False)
    pg.adjust_positions(0, True)
    po.adjust_positions(0, True)
    for k in range(len(truth)):
        assert pg.get_position(k) == po.get_position(k) == truth[k]
    # sub-pixel granularity (interpolated frame and support): device = oracle
    pg, po, ests, xs, keep = build(nd, oracle, d, starts, granularity=2)
    pg.adjust_positions(0, True)
    po.adjust_positions(0, True)
    for k in range(len(truth)):
        assert pg.get_position(k) == po.get_position(k)


def test_adjust_positions_rejects_bad_type(nd, oracle):
    d = synth.ptycho_frames()
    pg, po, ests, xs, keep = build(nd, oracle, d, d["positions"])
    with pytest.raises(Exception):
        pg.adjust_positions(5)
