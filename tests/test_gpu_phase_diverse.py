"""PhaseDiverseCDI on the device (nadia_pd) against the oracle: series and
parallel modes, planar and Fresnel frames, sub-pixel granularity."""
import numpy as np
import pytest

from test_gpu_fresnel import fresnel_pair

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def planar_frames(nd, oracle, synth, positions, n=64, alg="ER", double=True):
    gpu, cpu = [], []
    for k, _ in enumerate(positions):
        d = synth.planar_block(n=n, block=24, support=30, seed_amp=10 + k, seed_poisson=10 + k, seed_phase=20 + k)
        cdt = np.complex128 if double else np.complex64
        est = nd.Complex2D(n, n, double=double, data=d["x0"].astype(cdt))
        g = nd.PlanarCDI(est)
        xo = d["x0"].astype(np.complex128 if double else np.complex64).copy()
        o = oracle.OracleCDI(xo, oracle.PLANAR, double=double)
        for obj in (g, o):
            obj.set_support(d["support"])
            obj.set_intensity(d["intensity"])
            obj.set_algorithm(getattr(oracle, alg))
        gpu.append((est, g))
        cpu.append((xo, o))
    return gpu, cpu


@pytest.mark.parametrize("parallel", [False, True], ids=["series", "parallel"])
@pytest.mark.parametrize("scale", [1, 2])
def test_planar_frames(nd, oracle, synth, parallel, scale):
    pos = [(0, 0), (12.5, -7.25), (-20, 9)]
    gpu, cpu = planar_frames(nd, oracle, synth, pos, alg="HIO")
    pg = nd.PhaseDiverseCDI(beta=0.8, gamma=1.0, parallel=parallel, granularity=scale)
    po = oracle.OraclePhaseDiverse(beta=0.8, gamma=1.0, parallel=parallel, granularity=scale)
    for (est, g), (xo, o), (x, y) in zip(gpu, cpu, pos):
        pg.add_new_position(g, x, y, 1.0)
        po.add_new_position(o, x, y, 1.0)
    assert pg.size() == po.size()
    pg.set_iterations_per_cycle(2)
    po.set_iterations_per_cycle(2)
    pg.initialise_estimate()
    po.initialise_estimate()
    assert rel(pg.get_transmission(), po.get_transmission()) <= 1e-14
    for it in range(3):
        eg = pg.iterate()
        po.iterate()
        eo = np.array([o.get_error() for _, o in cpu])
        assert (np.abs(eg - eo) / eo).max() <= 1e-9, it
        assert rel(pg.get_transmission(), po.get_transmission()) <= 1e-10, it
    for (est, g), (xo, o) in zip(gpu, cpu):
        assert rel(est.get(), xo) <= 1e-10


@pytest.mark.parametrize("parallel", [False, True], ids=["series", "parallel"])
def test_fresnel_frames(nd, oracle, synth, parallel):
    pg = nd.PhaseDiverseCDI(beta=0.5, gamma=1.5, parallel=parallel)
    po = oracle.OraclePhaseDiverse(beta=0.5, gamma=1.5, parallel=parallel)
    keep = []
    for k, (x, y) in enumerate([(0, 0), (6, -4)]):
        d, est, g, xo, o = fresnel_pair(nd, oracle, synth)
        g.initialise_estimate(k)
        o.initialise_estimate(k)
        pg.add_new_position(g, x, y, 1.0 + 0.5 * k)
        po.add_new_position(o, x, y, 1.0 + 0.5 * k)
        keep.append((est, g, xo, o))
    pg.initialise_estimate()
    po.initialise_estimate()
    assert rel(pg.get_transmission(), po.get_transmission()) <= 1e-13
    for it in range(3):
        pg.iterate()
        po.iterate()
    assert rel(pg.get_transmission(), po.get_transmission()) <= 1e-10
    for est, g, xo, o in keep:
        assert rel(est.get(), xo) <= 1e-10


def test_fp32_frames(nd, oracle, synth):
    pos = [(0, 0), (8, 8)]
    gpu, cpu = planar_frames(nd, oracle, synth, pos, alg="ER", double=False)
    pg = nd.PhaseDiverseCDI(beta=1.0, parallel=False)
    po = oracle.OraclePhaseDiverse(beta=1.0, parallel=False, double=False)
    for (est, g), (xo, o), (x, y) in zip(gpu, cpu, pos):
        pg.add_new_position(g, x, y)
        po.add_new_position(o, x, y)
    pg.initialise_estimate()
    po.initialise_estimate()
    for _ in range(4):
        pg.iterate()
        po.iterate()
    assert rel(pg.get_transmission(), po.get_transmission()) <= 1e-5


@pytest.mark.parametrize("step_y", [7.5, 7.0], ids=["fractional", "integral"])
def test_many_frames(nd, oracle, synth, step_y):
    """More frames than one update launch, one object pass and one error
    gather take (16): 18 planar frames on a 6 x 3 grid of positions, parallel
    mode; integral positions take the one-pass add of all frames."""
    pos = [(9.0 * (k % 6) - 20, step_y * (k // 6) - 8) for k in range(18)]
    gpu, cpu = planar_frames(nd, oracle, synth, pos, n=32, alg="HIO")
    pg = nd.PhaseDiverseCDI(beta=0.9, gamma=1.0, parallel=True)
    po = oracle.OraclePhaseDiverse(beta=0.9, gamma=1.0, parallel=True)
    for (est, g), (xo, o), (x, y) in zip(gpu, cpu, pos):
        pg.add_new_position(g, x, y, 1.0)
        po.add_new_position(o, x, y, 1.0)
    pg.initialise_estimate()
    po.initialise_estimate()
    for it in range(3):
        eg = pg.iterate()
        po.iterate()
        eo = np.array([o.get_error() for _, o in cpu])
        assert eg.shape == (18,)
        assert (np.abs(eg - eo) / eo).max() <= 1e-9, it
    assert rel(pg.get_transmission(), po.get_transmission()) <= 1e-10


def test_fresnel_frames_with_constraint(nd, oracle, synth):
    """Frames carrying a TransmissionConstraint: get_result applies it before
    add_to_object (the two steps stay apart), the frames iterate on the
    unfused path."""
    pg = nd.PhaseDiverseCDI(beta=0.7, gamma=1.0, parallel=True)
    po = oracle.OraclePhaseDiverse(beta=0.7, gamma=1.0, parallel=True)
    keep = []
    rng = np.random.default_rng(7)
    for k, (x, y) in enumerate([(0, 0), (5, -3)]):
        d, est, g, xo, o = fresnel_pair(nd, oracle, synth)
        n = est.get().shape[0]
        tg = nd.TransmissionConstraint(n, n)
        tcs = oracle.OracleConstraint(n, n)
        for tc in (tg, tcs):
            tc.add_complex_constraint(d["support"], 0.5, 0.5)
        g.set_complex_constraint(tg)
        o.set_complex_constraint(tcs)
        # phases away from zero keep the region constant well conditioned
        T0 = np.exp((0.3 + 1j) * -rng.uniform(0.1, 0.4, (n, n)))
        g.set_transmission_function(T0)
        o.set_transmission_function(T0)
        pg.add_new_position(g, x, y, 1.0)
        po.add_new_position(o, x, y, 1.0)
        keep.append((est, g, xo, o, tg, tcs))
    pg.initialise_estimate()
    po.initialise_estimate()
    assert rel(pg.get_transmission(), po.get_transmission()) <= 1e-12
    for it in range(3):
        pg.iterate()
        po.iterate()
        assert rel(pg.get_transmission(), po.get_transmission()) <= 1e-10, it
    for est, g, xo, o, tg, tcs in keep:
        assert rel(est.get(), xo) <= 1e-10


@pytest.mark.parametrize("parallel", [False, True], ids=["series", "parallel"])
@pytest.mark.parametrize("scale", [1, 2])
def test_rectangular_planar_frames(nd, oracle, parallel, scale):
    """Frames of 48 x 96 pixels at integral and fractional positions: the
    object's bounding box, the weights and the frame <-> object mappings per
    axis, against the oracle (errors 1e-9, object and frames 1e-10)."""
    shape = (48, 96)
    nx, ny = shape
    pos = [(0, 0), (10.5, -13.25), (-7, 21)]
    gpu, cpu = [], []
    for k, _ in enumerate(pos):
        rng = np.random.default_rng(40 + k)
        obj = np.zeros(shape, complex)
        obj[nx // 4:3 * nx // 4, ny // 4:3 * ny // 4] = rng.uniform(0.5, 1, (nx // 2, ny // 2))
        I = np.abs(np.fft.fft2(obj)) ** 2 + 1e-3
        sup = np.zeros(shape)
        sup[nx // 8:7 * nx // 8, ny // 8:7 * ny // 8] = 1.0
        x0 = sup * np.exp(2j * np.pi * rng.random(shape))
        est = nd.Complex2D(nx, ny, double=True, data=x0)
        g = nd.PlanarCDI(est)
        xo = x0.copy()
        o = oracle.OracleCDI(xo, oracle.PLANAR)
        for obj_ in (g, o):
            obj_.set_support(sup)
            obj_.set_intensity(I)
            obj_.set_algorithm(oracle.HIO)
        gpu.append((est, g))
        cpu.append((xo, o))
    pg = nd.PhaseDiverseCDI(beta=0.8, gamma=1.0, parallel=parallel, granularity=scale)
    po = oracle.OraclePhaseDiverse(beta=0.8, gamma=1.0, parallel=parallel, granularity=scale)
    for (est, g), (xo, o), (x, y) in zip(gpu, cpu, pos):
        pg.add_new_position(g, x, y, 1.0)
        po.add_new_position(o, x, y, 1.0)
    assert pg.size() == po.size()
    pg.set_iterations_per_cycle(2)
    po.set_iterations_per_cycle(2)
    pg.initialise_estimate()
    po.initialise_estimate()
    assert rel(pg.get_transmission(), po.get_transmission()) <= 1e-14
    for it in range(3):
        eg = pg.iterate()
        po.iterate()
        eo = np.array([o.get_error() for _, o in cpu])
        assert (np.abs(eg - eo) / eo).max() <= 1e-9, it
        assert rel(pg.get_transmission(), po.get_transmission()) <= 1e-10, it
    for (est, g), (xo, o) in zip(gpu, cpu):
        assert rel(est.get(), xo) <= 1e-10
