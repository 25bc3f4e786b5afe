"""Transforms of any length and the unfused iteration at sizes without
compiled single-launch passes.

FFTW behind Complex_2D (NADIA/src/Complex_2D.c++:472-505) plans any nx x ny;
here lengths other than 2^k / 3*2^k run kern_any.inc: direct mixed-radix
plans (radices 2..13 in FP32, 2..7 in FP64) or Bluestein's algorithm.
Checked against numpy.fft (an independent FFT: pocketfft), against the
oracle's own mixed-radix FFT, and reconstructions against the oracle at
those sizes (FP64 one iteration <= 1e-12, trajectories <= 1e-9)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# (shape, what the two axes exercise)
FFT_SHAPES = [
    (10, 14),      # 2*5, 2*7: direct
    (100, 126),    # 4*5*5, 2*3*3*7
    (74, 130),     # 2*37 Bluestein; 2*5*13 direct in FP32, Bluestein in FP64
    (17, 34),      # odd prime rows of Bluestein length 64
    (1, 8),        # a length-1 axis is the identity
    (1000, 1000),  # 8*5*5*5
    (1030, 64),    # 2*5*103: Bluestein of length 4096 along the columns
    (4095, 6),     # 3*3*5*7*13: FP64 Bluestein of length 8192 (the largest LDS tile)
    (6, 4093),     # prime rows of Bluestein length 8192
    (48, 40),      # a native length on one axis only
]


@pytest.mark.parametrize("double", [True, False], ids=["fp64", "fp32"])
@pytest.mark.parametrize("shape", FFT_SHAPES, ids=lambda s: "%dx%d" % s)
def test_fft_any_length_matches_numpy(nd, shape, double):
    rng = np.random.default_rng(sum(shape))
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    cdt = np.complex128 if double else np.complex64
    a = a.astype(cdt)
    fwd = np.fft.fft2(a.astype(np.complex128))
    bwd = np.fft.ifft2(a.astype(np.complex128)) * a.size
    tol = 1e-12 if double else 2e-6
    f = nd.Complex2D(*shape, double=double, data=a)
    f.perform_forward_fft()
    assert rel(f.get(), fwd) <= tol, "forward"
    f.set(a)
    f.perform_backward_fft()
    assert rel(f.get(), bwd) <= tol, "backward"
    # forward then backward: n * identity
    f.set(a)
    f.perform_forward_fft()
    f.perform_backward_fft()
    assert rel(f.get() / a.size, a) <= tol


@pytest.mark.parametrize("shape", [(100, 126), (74, 130), (60, 22)], ids=lambda s: "%dx%d" % s)
def test_fft_any_length_matches_oracle(nd, oracle, shape):
    rng = np.random.default_rng(5)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    for sign, fn in ((-1, "perform_forward_fft"), (+1, "perform_backward_fft")):
        f = nd.Complex2D(*shape, double=True, data=a)
        getattr(f, fn)()
        b = a.copy()
        oracle.fft2d(b, sign)
        assert rel(f.get(), b) <= 1e-12, sign


def test_size_queries(nd):
    lib = nd.library()
    for n, native in ((2048, 1), (384, 1), (100, 0), (1030, 0), (1, 0), (4096, 1), (4093, 0)):
        assert lib.nadia_size_supported(n) == 1
        assert lib.nadia_size_native(n) == native
    for n in (0, 4097, -3):
        assert lib.nadia_size_supported(n) == 0


def planar_state(shape, seed=11):
    nx, ny = shape
    rng = np.random.default_rng(seed)
    obj = np.zeros(shape)
    bx, by = nx // 3, ny // 3
    obj[nx // 2 - bx // 2:nx // 2 - bx // 2 + bx, ny // 2 - by // 2:ny // 2 - by // 2 + by] = rng.uniform(0.5, 1, (bx, by))
    inten = np.abs(np.fft.fftshift(np.fft.fft2(obj))) ** 2
    inten = rng.poisson(inten * 1e6 / inten.max()).astype(np.float64)
    sup = np.zeros(shape)
    sx, sy = nx // 2, ny // 2
    sup[nx // 2 - sx // 2:nx // 2 - sx // 2 + sx, ny // 2 - sy // 2:ny // 2 - sy // 2 + sy] = 1.0
    x0 = sup * np.exp(1j * rng.uniform(0, 2 * np.pi, shape))
    return dict(intensity=inten, support=sup, x0=x0)


PLANAR_SHAPES = [(100, 126), (74, 130), (60, 22)]


@pytest.mark.parametrize("path", [0, 1], ids=["fused", "plain"])
@pytest.mark.parametrize("shape", PLANAR_SHAPES, ids=lambda s: "%dx%d" % s)
@pytest.mark.parametrize("alg", ["HIO", "RAAR", "ER", "DM"])
def test_planar_any_size_matches_oracle(nd, oracle, shape, alg, path):
    """Fused: kern_any.inc's column and row passes (DM, with its P_F P_S
    term, runs unfused either way); plain: the reference's data flow."""
    d = planar_state(shape)
    est = nd.Complex2D(*shape, data=d["x0"])
    g = nd.PlanarCDI(est)
    g.set_path(path)
    xo = d["x0"].copy()
    o = oracle.OracleCDI(xo, oracle.PLANAR)
    a = nd.get_alg_from_name(alg)
    for obj in (g, o):
        obj.set_support(d["support"])
        obj.set_intensity(d["intensity"])
        obj.set_algorithm(a)
    e = g.iterate(1)[0]
    o.iterate()
    assert rel(est.get(), xo) <= 1e-12
    assert abs(e - o.get_error()) <= 1e-12 * o.get_error()
    eg = g.iterate(30)
    eo = o.iterate_n(30)
    assert (np.abs(eg - eo) / eo).max() <= 1e-9
    # switching paths mid-run continues from the same state
    g.set_path(1 - path)
    eg = g.iterate(5)
    eo = o.iterate_n(5)
    assert (np.abs(eg - eo) / eo).max() <= 1e-9


def test_planar_any_size_shrinkwrap_and_fp32(nd, oracle):
    shape = (100, 126)
    d = planar_state(shape, seed=12)
    est = nd.Complex2D(*shape, data=d["x0"])
    g = nd.PlanarCDI(est)
    xo = d["x0"].copy()
    o = oracle.OracleCDI(xo, oracle.PLANAR)
    for obj in (g, o):
        obj.set_support(d["support"])
        obj.set_intensity(d["intensity"])
        obj.set_algorithm(nd.HIO)
    g.iterate(20)
    o.iterate_n(20)
    g.apply_shrinkwrap(2.0, 0.1)
    o.apply_shrinkwrap(2.0, 0.1)
    # BaseCDI::apply_threshold loops j < nx (NADIA/src/BaseCDI.c++:477-492):
    # on a 100 x 126 field the columns j >= 100 keep the convolved magnitude,
    # the rest is thresholded to 0 / 1, and set_support divides everything
    # by the maximum, which the soft columns set.  The 0 / 1 pattern must
    # match exactly; the values carry the estimate's last-bit differences.
    sg, so = g.get_support(), o.get_support()
    assert np.array_equal(sg[:, :100] > 0, so[:, :100] > 0)
    assert len(np.unique(sg[:, :100])) == 2
    assert rel(sg, so) <= 1e-10
    assert sg[:, 100:].max() == 1.0
    eg = g.iterate(10)
    eo = o.iterate_n(10)
    assert (np.abs(eg - eo) / eo).max() <= 1e-9
    # FP32 bulk against the FP64 oracle
    d = planar_state(shape, seed=13)
    est = nd.Complex2D(*shape, double=False, data=d["x0"].astype(np.complex64))
    g = nd.PlanarCDI(est)
    xo = d["x0"].copy()
    o = oracle.OracleCDI(xo, oracle.PLANAR)
    for obj in (g, o):
        obj.set_support(d["support"])
        obj.set_intensity(d["intensity"])
        obj.set_algorithm(nd.HIO)
    eg = g.iterate(5)
    eo = o.iterate_n(5)
    assert rel(est.get().astype(np.complex128), xo) <= 1e-4
    assert (np.abs(eg - eo) / eo).max() <= 1e-4


@pytest.mark.parametrize("shape", [(100, 126), (74, 130)], ids=lambda s: "%dx%d" % s)
def test_fused_any_size_fp32_and_n_best(nd, oracle, shape):
    """FP32 bulk on the any-length fused passes against the FP64 oracle, and
    the best estimates kept on the device."""
    d = planar_state(shape, seed=15)
    est = nd.Complex2D(*shape, double=False, data=d["x0"].astype(np.complex64))
    g = nd.PlanarCDI(est, n_best=3)
    xo = d["x0"].copy()
    o = oracle.OracleCDI(xo, oracle.PLANAR, n_best=3)
    for obj in (g, o):
        obj.set_support(d["support"])
        obj.set_intensity(d["intensity"])
        obj.set_algorithm(nd.HIO)
    eg = g.iterate(6)
    eo = o.iterate_n(6)
    assert rel(est.get().astype(np.complex128), xo) <= 1e-4
    assert (np.abs(eg - eo) / eo).max() <= 1e-4
    for k in range(3):
        bg, e_g = g.get_best_result(k)
        bo, e_o = o.get_best_result(k)
        assert abs(e_g - e_o) <= 1e-4 * e_o
        assert rel(bg.astype(np.complex128), bo) <= 1e-4


def test_shrinkwrap_tall_field(nd, oracle):
    """nx > ny: every pixel thresholded once (the reference's loop over
    j < nx overruns the rows there)."""
    shape = (126, 100)
    d = planar_state(shape, seed=14)
    est = nd.Complex2D(*shape, data=d["x0"])
    g = nd.PlanarCDI(est)
    xo = d["x0"].copy()
    o = oracle.OracleCDI(xo, oracle.PLANAR)
    for obj in (g, o):
        obj.set_support(d["support"])
        obj.set_intensity(d["intensity"])
    g.apply_shrinkwrap(1.5, 0.1)
    o.apply_shrinkwrap(1.5, 0.1)
    sg, so = g.get_support(), o.get_support()
    assert np.array_equal(sg, so)
    assert set(np.unique(sg)) == {0.0, 1.0}


@pytest.mark.parametrize("path", [0, 1], ids=["fused", "plain"])
def test_fresnel_any_size_matches_oracle(nd, oracle, path):
    nx, ny = 100, 126
    rng = np.random.default_rng(21)
    wf = np.exp(1j * rng.uniform(0, 0.3, (nx, ny))) * (1 + 0.1 * rng.uniform(size=(nx, ny)))
    inten = rng.uniform(0.5, 1.5, (nx, ny)) * np.abs(wf) ** 2
    sup = np.zeros((nx, ny))
    sup[25:75, 30:96] = 1.0
    par = dict(wavelength=1.55e-10, zfd=0.909, zfs=1.6e-3, pixel=13.5e-6)
    est = nd.Complex2D(nx, ny)
    g = nd.FresnelCDI(est, wf, par["wavelength"], par["zfd"], par["zfs"], par["pixel"], 1.0)
    g.set_path(path)
    xo = np.zeros((nx, ny), np.complex128)
    o = oracle.OracleCDI(xo, oracle.FRESNEL)
    o.fresnel_init(wf, par["wavelength"], par["zfd"], par["zfs"], par["pixel"], 1.0)
    for obj in (g, o):
        obj.set_support(sup)
        obj.set_intensity(inten)
    g.initialise_estimate(0)
    o.initialise_estimate(0)
    assert rel(est.get(), xo) <= 1e-14
    for alg in ("ER", "HIO"):
        a = nd.get_alg_from_name(alg)
        g.set_algorithm(a)
        o.set_algorithm(a)
        e = g.iterate(1)[0]
        o.iterate()
        assert rel(est.get(), xo) <= 1e-12, alg
        assert abs(e - o.get_error()) <= 1e-12 * o.get_error(), alg
    eg = g.iterate(20)
    eo = o.iterate_n(20)
    assert (np.abs(eg - eo) / eo).max() <= 1e-9


def test_wf_any_size_matches_oracle(nd, oracle):
    nx, ny = 100, 126
    rng = np.random.default_rng(22)
    inten = rng.uniform(0.5, 1.5, (nx, ny))
    par = dict(wavelength=1.55e-10, z12=16.353e-3, z23=0.909, pixel=13.5e-6)
    est = nd.Complex2D(nx, ny)
    g = nd.FresnelCDI_WF(est, par["wavelength"], par["z12"], par["z23"], par["pixel"])
    xo = np.zeros((nx, ny), np.complex128)
    o = oracle.OracleCDI(xo, oracle.FRESNEL_WF)
    o.wf_init(par["wavelength"], par["z12"], par["z23"], par["pixel"])
    g.set_support(40e-6)
    o.wf_set_support(40e-6)
    assert rel(g.get_support(), o.get_support()) <= 1e-14
    for obj in (g, o):
        obj.set_intensity(inten)
    g.initialise_estimate(0)
    o.initialise_estimate(0)
    e = g.iterate(1)[0]
    o.iterate()
    assert rel(est.get(), xo) <= 1e-12
    assert abs(e - o.get_error()) <= 1e-12 * o.get_error()
    eg = g.iterate(20)
    eo = o.iterate_n(20)
    assert (np.abs(eg - eo) / eo).max() <= 1e-9


def test_odd_reconstruction_sizes_refused(nd):
    f = nd.Complex2D(99, 126)
    r = ctypes.c_void_p()
    assert nd.library().nadia_recon_create(f.h, nd.PLANAR, 0, ctypes.byref(r)) == -3
    assert "odd" in nd.library().nadia_last_error().decode()


def test_batch_mixes_native_and_any_sizes(nd):
    """A batch runs the any-size jobs unfused beside the shared launches of
    the native ones; every job equals the same job run alone."""
    shapes = [(100, 126), (128, 128), (74, 130), (128, 128)]
    runs = []
    for alone in (True, False):
        jobs = []
        for k, sh in enumerate(shapes):
            d = planar_state(sh, seed=30 + k)
            est = nd.Complex2D(*sh, double=False, data=d["x0"].astype(np.complex64))
            g = nd.PlanarCDI(est)
            g.set_support(d["support"])
            g.set_intensity(d["intensity"])
            g.set_algorithm(nd.RAAR)
            jobs.append((est, g))
        if alone:
            for _, g in jobs:
                g.iterate(12)
        else:
            b = nd.Batch([g for _, g in jobs])
            b.iterate(12)
        runs.append([est.get() for est, _ in jobs])
    for a, b in zip(*runs):
        assert np.array_equal(a, b)
