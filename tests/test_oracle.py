"""Pinning of the CPU oracle (oracle/nadia_oracle.c) without the reference.

Each check is chosen so that a plausible mistake (wrong sign, dropped term,
transposed index, off-by-one centre) fails: brute-force DFT, numpy.fft as an
independent library, DFT closed forms, fftshift identity, unitarity and round
trips, the named-algorithm coefficient table, fixed points of every
algorithm, HIO's inside/outside rule, scipy's 2-D convolution for the
shrinkwrap window, closed-form Fresnel / white-field phases and the C
library's rand() sequence for initialise_estimate.
"""
import ctypes

import numpy as np
import pytest


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def brute_dft2(a, sign):
    nx, ny = a.shape
    i = np.arange(nx)
    j = np.arange(ny)
    # exponents reduced exactly in integers before scaling
    Wx = np.exp(sign * 2j * np.pi * (np.outer(i, i) % nx) / nx)
    Wy = np.exp(sign * 2j * np.pi * (np.outer(j, j) % ny) / ny)
    return Wx @ a @ Wy


# ------------------------------------------------------------------ FFT ----

@pytest.mark.parametrize("shape", [(8, 8), (4, 16), (12, 24), (24, 8), (6, 10), (32, 32), (48, 16)])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft_brute_force(oracle, shape, sign):
    rng = np.random.default_rng(0)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    b = a.copy()
    oracle.fft2d(b, sign)
    assert rel(b, brute_dft2(a, sign)) < 1e-14


@pytest.mark.parametrize("shape", [(64, 64), (256, 128), (384, 384), (768, 96), (2048, 16)])
def test_fft_numpy(oracle, shape):
    rng = np.random.default_rng(1)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    b = a.copy()
    oracle.fft2d(b, -1)
    assert rel(b, np.fft.fft2(a)) < 1e-14
    c = a.copy()
    oracle.fft2d(c, +1)
    assert rel(c, np.fft.ifft2(a) * a.size) < 1e-14


def test_fft_closed_forms(oracle):
    nx, ny = 32, 48
    d = np.zeros((nx, ny), complex)
    d[0, 0] = 1
    oracle.fft2d(d, -1)
    assert np.allclose(d, 1, atol=1e-15)  # delta -> constant
    k0, l0 = 5, 7
    x = np.arange(nx)[:, None]
    y = np.arange(ny)[None, :]
    pw = np.exp(2j * np.pi * (k0 * x / nx + l0 * y / ny))
    oracle.fft2d(pw, -1)  # plane wave -> N delta at (k0, l0)
    expect = np.zeros((nx, ny), complex)
    expect[k0, l0] = nx * ny
    assert np.abs(pw - expect).max() < 1e-11
    rng = np.random.default_rng(2)
    a = rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))
    b = a.copy()
    oracle.fft2d(b, -1)  # Parseval
    assert abs(np.sum(np.abs(b) ** 2) / (nx * ny) - np.sum(np.abs(a) ** 2)) < 1e-10 * np.sum(np.abs(a) ** 2)


def test_fft_fp32_variant(oracle):
    rng = np.random.default_rng(3)
    a = (rng.standard_normal((128, 96)) + 1j * rng.standard_normal((128, 96))).astype(np.complex64)
    b = a.copy()
    oracle.fft2d(b, -1, double=False)
    assert rel(b, np.fft.fft2(a.astype(np.complex128))) < 1e-6


def test_invert_is_scaled_fftshift(oracle):
    rng = np.random.default_rng(4)
    a = rng.standard_normal((16, 24)) + 1j * rng.standard_normal((16, 24))
    b = a.copy()
    oracle.invert(b, True)
    assert rel(b, np.fft.fftshift(a) / np.sqrt(a.size)) < 1e-15
    c = a.copy()
    oracle.invert(c, False)
    oracle.invert(c, False)
    assert np.array_equal(c, a)  # an involution without scaling


# ------------------------------------------------------ planar operators ----

def planar(oracle, n=32, seed=5, beam=False, kind=None, x=None):
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) if x is None else x
    o = oracle.OracleCDI(x, oracle.PLANAR if kind is None else kind)
    I = rng.uniform(0.5, 4.0, (n, n)) ** 2
    o.set_intensity(I)
    sup = np.zeros((n, n))
    sup[n // 4:3 * n // 4, n // 4:3 * n // 4] = 1
    o.set_support(sup)
    bs = None
    if beam:
        bs = np.ones((n, n))
        bs[n // 2 - 2:n // 2 + 2, :] = 0
        bs[0, 0] = -1.0
        o.set_beam_stop(bs)
    return o, x, I, sup, bs


def test_planar_propagators_unitary_and_centred(oracle):
    o, x, *_ = planar(oracle, n=32)
    y = x.copy()
    o.propagate_to_detector(y)
    assert abs(np.linalg.norm(y) - np.linalg.norm(x)) < 1e-13 * np.linalg.norm(x)
    assert rel(y, np.fft.fftshift(np.fft.fft2(x)) / 32)  < 1e-14  # centred unitary DFT
    o.propagate_from_detector(y)
    assert rel(y, x) < 1e-14
    c = np.ones((32, 32), complex)
    o.propagate_to_detector(c)
    assert abs(c[16, 16] - 32) < 1e-12 and np.abs(c).sum() - abs(c[16, 16]) < 1e-10  # DC at the centre


def test_scale_intensity_properties(oracle):
    o, x, I, sup, bs = planar(oracle, n=32, beam=True)
    y = x.copy()
    y[3, 4] = 0
    o.scale_intensity(y)
    q = np.sqrt(I)
    proj = bs > 0
    assert np.allclose(np.abs(y[proj]), q[proj], rtol=1e-15)  # modulus replaced
    keep = proj.copy()
    keep[3, 4] = False  # zeroed above
    assert np.allclose(y[keep] / np.abs(y[keep]), x[keep] / np.abs(x[keep]), atol=1e-15)  # phase kept
    assert np.array_equal(y[~proj], x[~proj])  # beam-stop pixels (value <= 0) free
    assert y[3, 4] == q[3, 4]  # |y| = 0 -> (sqrt I, 0)
    # error: zero for the exact modulus, 1 for twice the modulus
    t = q * np.exp(1j * np.angle(x))
    o.scale_intensity(t.copy())
    assert o.get_error() < 1e-30
    o2, *_ = planar(oracle, n=32, seed=5)
    o2.scale_intensity(2 * t)
    assert abs(o2.get_error() - 1.0) < 1e-14


COEF = {  # (PSF, PFS, PS, PF, PI) of the named algorithms, beta = 0.9
    "BIO": (0.9, 0, 0, -0.9, 0),
    "BOO": (0.9, 0, 0, 0.1, -1),
    "HIO": (1.9, 0, -1, -0.9, 0),
    "DM": (1.9, 0.1, -1, -1, 0),
    "SF": (2, 0, 0, -1, -1),
    "ASR": (2, 0, -1, -1, 0),
    "HPR": (1.9, 0, -1, -0.9, 0),
    "RAAR": (1.9, 0, -0.9, -0.9, -0.1),
}
ALG = {"ER": 0, "BIO": 1, "BOO": 2, "HIO": 3, "DM": 4, "SF": 5, "ASR": 6, "HPR": 7, "RAAR": 8}


@pytest.mark.parametrize("name", sorted(COEF))
def test_coefficient_table(oracle, name):
    o, *_ = planar(oracle)
    o.set_algorithm(ALG[name])
    c, alg = o.coefficients()
    assert alg == ALG[name]
    assert np.allclose(c, COEF[name], atol=1e-15)
    assert abs(c.sum()) < 1e-15


def test_custom_coefficients_sum_to_zero(oracle):
    rng = np.random.default_rng(6)
    o, *_ = planar(oracle)
    for _ in range(5):
        o.set_custom_algorithm(*rng.uniform(-2, 2, 10))
        c, alg = o.coefficients()
        assert alg == 9 and abs(c.sum()) < 1e-14


def consistent(oracle, n=32):
    """x supported inside the support and I = |F x|^2: every algorithm's fixed point."""
    rng = np.random.default_rng(7)
    x = np.zeros((n, n), complex)
    x[n // 4:3 * n // 4, n // 4:3 * n // 4] = rng.standard_normal((n // 2, n // 2)) + 1j
    I = np.abs(np.fft.fftshift(np.fft.fft2(x)) / n) ** 2
    sup = np.zeros((n, n))
    sup[n // 4 - 2:3 * n // 4 + 2, n // 4 - 2:3 * n // 4 + 2] = 1
    return x, I, sup


@pytest.mark.parametrize("name", ["ER"] + sorted(COEF))
def test_fixed_point_every_algorithm(oracle, name):
    x, I, sup = consistent(oracle)
    xe = x.copy()
    o = oracle.OracleCDI(xe)
    o.set_support(sup)
    o.set_intensity(I)
    o.set_algorithm(ALG[name])
    o.iterate()
    assert rel(xe, x) < 1e-13
    assert o.get_error() < 1e-26


def test_hio_inside_outside_rule_and_er(oracle):
    o, x, I, sup, _ = planar(oracle, n=32, seed=8)
    pf = x.copy()
    o.project_intensity(pf)
    xe = x.copy()
    o2 = oracle.OracleCDI(xe)
    o2.set_intensity(I)
    o2.set_support(sup)
    o2.set_algorithm(ALG["HIO"])
    o2.iterate()
    expect = np.where(sup == 1, pf, x - 0.9 * pf)
    assert rel(xe, expect) < 1e-14
    xe2 = x.copy()
    o3 = oracle.OracleCDI(xe2)
    o3.set_intensity(I)
    o3.set_support(sup)
    o3.iterate()  # ER by default (BaseCDI constructor)
    assert rel(xe2, pf * sup) < 1e-14


def test_dm_uses_two_projections(oracle):
    """DM: (1+b) Ps Pf x + (1-b) Pf Ps x - Ps x - Pf x + x; get_error is the
    error of the last projection (P_F x)."""
    o, x, I, sup, _ = planar(oracle, n=32, seed=9)
    pf = x.copy()
    o.project_intensity(pf)
    e_pf = o.get_error()
    pfs = x * sup
    o.project_intensity(pfs)
    xe = x.copy()
    o2 = oracle.OracleCDI(xe)
    o2.set_intensity(I)
    o2.set_support(sup)
    o2.set_algorithm(ALG["DM"])
    o2.iterate()
    expect = x + 0.1 * pfs - pf - x * sup + 1.9 * pf * sup
    assert rel(xe, expect) < 1e-14
    assert abs(o2.get_error() - e_pf) < 1e-15 * e_pf


def test_relaxation_quirk(oracle):
    o, *_ = planar(oracle)
    o.set_algorithm(ALG["HIO"])
    o.set_relaxation_parameter(0.5)
    assert np.allclose(o.coefficients()[0], COEF["HIO"])  # unchanged
    o.set_relaxation_parameter(0.5)
    o.set_algorithm(ALG["RAAR"])
    assert np.allclose(o.coefficients()[0], (1.5, 0, -0.5, -0.5, -0.5))


def test_shrinkwrap_convolution_scipy(oracle):
    from scipy.signal import convolve2d

    o, x, *_ = planar(oracle, n=48, seed=10)
    a = np.abs(x)
    for sigma, cut in ((1.5, 4), (3.0, 5), (2.2, 4)):
        i = np.arange(-cut, cut + 1)
        g = np.exp(-(i[:, None] ** 2 + i[None, :] ** 2) / (2 * sigma * sigma))
        ref = convolve2d(a, g, mode="same", boundary="fill", fillvalue=0)
        assert rel(o.convolve(a, sigma, cut), ref) < 1e-14


def test_shrinkwrap_convolution_fp32_products(oracle):
    """Single-precision build: BaseCDI::convolve forms each product
    array.get() * gauss_dist.get() of two floats in float and adds it to a
    double (NADIA/src/BaseCDI.c++:455-458).  Restated as shift-and-add over
    the taps, di then dj ascending (the per-pixel order of the reference's
    loops), each tap's products rounded to float32; the same sum with the
    exact (double) products must differ, so the check tells the two apart."""
    rng = np.random.default_rng(21)
    n, m = 40, 36
    x = (rng.random((n, m)) + 1j * rng.random((n, m))).astype(np.complex64)
    of = oracle.OracleCDI(x.copy(), double=False)
    a = (rng.random((n, m)) * 3.0).astype(np.float32)
    for sigma, cut in ((1.5, 4), (3.0, 5)):
        i = np.arange(cut + 1)
        g = np.exp(-(i[:, None] ** 2 + i[None, :] ** 2) / (2.0 * sigma * sigma)).astype(np.float32)
        ap = np.zeros((n + 2 * cut, m + 2 * cut), np.float32)
        ap[cut:cut + n, cut:cut + m] = a
        acc_f = np.zeros((n, m))
        acc_d = np.zeros((n, m))
        for di in range(-cut, cut + 1):
            for dj in range(-cut, cut + 1):
                tap = ap[cut + di:cut + di + n, cut + dj:cut + dj + m]
                w = g[abs(di), abs(dj)]
                acc_f += (tap * w).astype(np.float64)  # float32 product
                acc_d += tap.astype(np.float64) * np.float64(w)
        got = of.convolve(a, sigma, cut)
        assert got.dtype == np.float32
        assert np.array_equal(got, acc_f.astype(np.float32))
        assert not np.array_equal(got, acc_d.astype(np.float32))


def test_shrinkwrap_threshold(oracle):
    from scipy.signal import convolve2d

    o, x, *_ = planar(oracle, n=48, seed=11)
    o.apply_shrinkwrap(2.0, 0.3)
    i = np.arange(-4, 5)
    g = np.exp(-(i[:, None] ** 2 + i[None, :] ** 2) / 8.0)
    c = convolve2d(np.abs(x), g, mode="same")
    expect = (c >= 0.3 * c.max()).astype(float)
    s = o.get_support()
    margin = np.abs(c - 0.3 * c.max()) > 1e-9 * c.max()  # ignore exact ties
    assert np.array_equal(s[margin], expect[margin])


def test_initialise_estimate_uses_libc_rand(oracle):
    """PlanarCDI::initialise_estimate: srand(seed), two rand() per support
    pixel in x-major order, scaled by int(max sqrt I) with the int product
    wrapping; zero outside the support."""
    libc = ctypes.CDLL("libc.so.6")
    n = 16
    rng = np.random.default_rng(12)
    x = np.zeros((n, n), complex)
    o = oracle.OracleCDI(x)
    I = rng.uniform(0, 50.0, (n, n)) ** 2
    I[0, 0] = 1e6 ** 2  # max sqrt I = 1e6: max_value * rand() overflows an int
    o.set_intensity(I)
    sup = (rng.uniform(size=(n, n)) > 0.5) * rng.uniform(0.2, 1.0, (n, n))
    o.set_support(sup)
    o.initialise_estimate(42)
    s = sup / sup.max()
    libc.srand(42)
    mv = int(np.sqrt(I).max())
    exp = np.zeros((n, n), complex)
    for i in range(n):
        for j in range(n):
            if s[i, j]:
                a = np.int32(np.uint32((mv * libc.rand()) & 0xFFFFFFFF)).item()
                b = np.int32(np.uint32((mv * libc.rand()) & 0xFFFFFFFF)).item()
                exp[i, j] = s[i, j] * (a / 2147483647.0) + 1j * s[i, j] * (b / 2147483647.0)
    assert rel(x, exp) < 1e-15
    assert np.all(x[s == 0] == 0)


def test_n_best_ordering(oracle):
    o, x, I, sup, _ = planar(oracle, n=32, seed=13)
    o2 = oracle.OracleCDI(x, n_best=3)
    o2.set_intensity(I)
    o2.set_support(sup)
    o2.set_algorithm(ALG["HIO"])
    errs = o2.iterate_n(8)
    best = sorted(errs)[:3]
    for k in range(3):
        _, e = o2.get_best_result(k)
        assert e == best[k]
    assert o2.get_best_result(3) == (None, None)


def test_fp32_oracle_close_to_fp64(oracle):
    o, x, I, sup, _ = planar(oracle, n=64, seed=14)
    xf = x.astype(np.complex64)
    of = oracle.OracleCDI(xf, double=False)
    of.set_intensity(I)
    of.set_support(sup)
    xd = x.copy()
    od = oracle.OracleCDI(xd)
    od.set_intensity(I)
    od.set_support(sup)
    for obj in (of, od):
        obj.set_algorithm(ALG["HIO"])
        obj.iterate()
    assert rel(xf.astype(np.complex128), xd) < 1e-5


# -------------------------------------------------------------- Fresnel ----

FR = dict(wavelength=1.55e-10, zfd=0.909, zfs=1.6e-3, pixel=13.5e-6)


def fresnel(oracle, n=32, seed=15):
    rng = np.random.default_rng(seed)
    wf = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 3.0
    x = np.zeros((n, n), complex)
    o = oracle.OracleCDI(x, oracle.FRESNEL)
    o.fresnel_init(wf, FR["wavelength"], FR["zfd"], FR["zfs"], FR["pixel"], 1.0)
    return o, x, wf, rng


def fresnel_phi(n):
    xm = (n - 1) / 2.0
    zsd = FR["zfd"] - FR["zfs"]
    f = FR["pixel"] ** 2 * np.pi / FR["wavelength"] * (1 / FR["zfd"] - 1 / zsd)
    i = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    return f * ((xm - i) ** 2 + (xm - j) ** 2)


def test_fresnel_propagator_closed_form_phase(oracle):
    n = 32
    o, x, wf, rng = fresnel(oracle, n)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    b = a.copy()
    o.propagate_to_detector(b)
    expect = np.fft.fft2(np.fft.fftshift(a) / n) * np.exp(-1j * fresnel_phi(n))
    assert rel(b, expect) < 1e-13  # mirrored coefficient table == closed form about (n-1)/2
    o.propagate_from_detector(b)
    assert rel(b, a) < 1e-13  # round trip


def test_fresnel_illumination_and_transmission(oracle):
    n = 32
    o, x, wf, rng = fresnel(oracle, n)
    Ls = o.get_illumination_at_sample()
    b = wf.copy()
    o.propagate_from_detector(b)
    assert rel(Ls, b) < 1e-15
    T = np.exp(-0.1 * rng.uniform(size=(n, n)) + 1j * rng.uniform(size=(n, n)))
    esw = np.zeros((n, n), complex)
    o.set_transmission_function(T, esw)
    assert rel(esw, (T - 1) * Ls) < 1e-14
    assert rel(o.get_transmission_function(esw), T) < 1e-13  # T -> ESW -> T


def test_fresnel_scale_intensity_adds_white_field(oracle):
    n = 32
    o, x, wf, rng = fresnel(oracle, n)
    I = rng.uniform(1, 9, (n, n))
    o.set_intensity(I)
    y = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    z = y.copy()
    o.scale_intensity(z)
    assert np.allclose(np.abs(z + wf), np.sqrt(I), rtol=1e-13)
    assert np.allclose(np.angle(z + wf), np.angle(y + wf), atol=1e-12)


def test_fresnel_norm(oracle):
    n = 32
    o, x, wf, rng = fresnel(oracle, n)
    I = rng.uniform(1, 9, (n, n))
    o.set_intensity(I)
    o.auto_set_norm()
    L = o.get_illumination()
    assert abs(np.abs(L).sum() - np.sqrt(I).sum()) < 1e-10 * np.sqrt(I).sum()


def test_fresnel_initialise_estimate(oracle):
    libc = ctypes.CDLL("libc.so.6")
    n = 16
    o, x, wf, rng = fresnel(oracle, n)
    sup = np.ones((n, n))
    o.set_support(sup)
    o.initialise_estimate(3)
    libc.srand(3)
    T = np.array([1 - 0.2 * libc.rand() / 2147483647.0 for _ in range(n * n)]).reshape(n, n)
    Ls = o.get_illumination_at_sample()
    assert rel(x, (T - 1) * Ls) < 1e-14


# ------------------------------------------------------ FresnelCDI_WF ----

WF = dict(wavelength=1.55e-10, z12=16.353e-3, z23=0.909, pixel=13.5e-6)


def wf_coef(n):
    xm = (n - 1) / 2.0
    s = WF["wavelength"] * WF["z23"] / (WF["pixel"] * n)
    i = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    rho2 = (s * (xm - i)) ** 2 + (s * (xm - j)) ** 2
    phi = (rho2 / WF["z12"] + rho2 / WF["z23"]) * np.pi / WF["wavelength"]
    return (np.sin(phi) - 1j * np.cos(phi)) / n


def test_wf_propagators_closed_form_and_round_trip(oracle):
    n = 32
    x = np.zeros((n, n), complex)
    o = oracle.OracleCDI(x, oracle.FRESNEL_WF)
    o.wf_init(WF["wavelength"], WF["z12"], WF["z23"], WF["pixel"])
    rng = np.random.default_rng(16)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    b = a.copy()
    o.propagate_from_detector(b)
    c = wf_coef(n)
    u = np.fft.fftshift(np.fft.ifft2(a) * n * n) / n * c
    assert rel(b, np.fft.ifft2(u) * n * n) < 1e-12
    o.propagate_to_detector(b)
    assert rel(b, a) < 1e-12


def test_wf_support_softened_disk(oracle):
    n = 32
    x = np.zeros((n, n), complex)
    o = oracle.OracleCDI(x, oracle.FRESNEL_WF)
    o.wf_init(WF["wavelength"], WF["z12"], WF["z23"], WF["pixel"])
    dz = WF["z12"] / WF["z23"] * WF["pixel"]
    o.wf_set_support(12 * dz, 1.0)
    s = o.get_support()
    c = (n - 1) // 2
    i = np.arange(n)[:, None] - c
    j = np.arange(n)[None, :] - c
    disk = (dz * np.sqrt(i * i + j * j) < 6 * dz).astype(float)
    from scipy.signal import convolve2d

    k = np.arange(-5, 6)
    g = np.exp(-(k[:, None] ** 2 + k[None, :] ** 2) / 18.0)
    assert rel(s, convolve2d(disk, g, mode="same")) < 1e-14
    assert s.max() > 1.0  # left unnormalised like the reference
    # (the white-field iteration itself, and its fixed point, are pinned in
    # test_oracle_pinning.py)


# ------------------------------------------------ TransmissionConstraint ----
# Pinned by what the mathematics fixes: a homogeneous material
# t = exp((kappa + i) phi), phi < 0, is a fixed point of every complex
# constraint and its region constant is c = kappa; alpha1 = 1 replaces the
# magnitude by exp(c phi); charge flipping negates positive imaginary parts;
# unity caps the magnitude at 1 keeping the phase; the first region to claim
# a pixel keeps it.

def homogeneous(n, kappa, rng):
    phi = -rng.uniform(0.05, 1.0, (n, n))
    return np.exp((kappa + 1j) * phi)


def test_constraint_homogeneous_fixed_point_and_c(oracle):
    n = 16
    rng = np.random.default_rng(40)
    for kappa in (0.3, 2.5):
        t = homogeneous(n, kappa, rng)
        tc = oracle.OracleConstraint(n, n)
        k = tc.add_complex_constraint(np.ones((n, n)), 0.7, 0.4)
        out = tc.apply(t.copy())
        assert abs(tc.c_mean(k) - kappa) < 1e-13 * kappa
        assert rel(out, t) < 1e-14


def test_constraint_alpha1_replaces_magnitude(oracle):
    n = 16
    rng = np.random.default_rng(41)
    kappa = 1.7
    phi = rng.uniform(0.05, 1.0, (n, n))        # positive phase: not a fixed point
    t = np.exp(-kappa * phi + 1j * phi)         # |log|t|| = kappa |phi|, so c = kappa
    tc = oracle.OracleConstraint(n, n)
    tc.set_charge_flipping(False)
    tc.set_enforce_unity(False)
    tc.add_complex_constraint(np.ones((n, n)), 1.0, 0.0)
    out = tc.apply(t.copy())
    assert rel(np.abs(out), np.exp(kappa * phi)) < 1e-13
    assert np.abs(np.angle(out) - phi).max() < 1e-13
    # alpha2 = 1 instead: phase log|t| / c = -phi, magnitude unchanged
    tc2 = oracle.OracleConstraint(n, n)
    tc2.set_charge_flipping(False)
    tc2.set_enforce_unity(False)
    tc2.add_complex_constraint(np.ones((n, n)), 0.0, 1.0)
    out2 = tc2.apply(t.copy())
    assert np.abs(np.angle(out2) + phi).max() < 1e-13
    assert rel(np.abs(out2), np.abs(t)) < 1e-13


def test_constraint_charge_flip_and_unity(oracle):
    n = 16
    rng = np.random.default_rng(42)
    t = (rng.uniform(0.2, 1.6, (n, n)) * np.exp(1j * rng.uniform(-np.pi, np.pi, (n, n))))
    tc = oracle.OracleConstraint(n, n)       # defaults: flip (sign +1) and unity
    out = tc.apply(t.copy())
    assert (out.imag <= 0).all()
    small = np.abs(t) <= 1
    assert np.array_equal(out[small], np.where(t.imag > 0, np.conj(t), t)[small])  # flip only: exact
    big = np.abs(t) > 1
    assert np.abs(np.abs(out[big]) - 1).max() < 1e-15
    assert np.abs(np.abs(out[~big]) - np.abs(t[~big])).max() < 1e-15
    flipped = np.where(t.imag > 0, np.conj(t), t)
    assert np.abs(np.angle(out) - np.angle(flipped)).max() < 1e-14
    tc2 = oracle.OracleConstraint(n, n)
    tc2.set_enforce_unity(False)
    tc2.set_charge_flipping(True, -1)         # flip negative imaginary parts instead
    out2 = tc2.apply(t.copy())
    assert (out2.imag >= 0).all() and rel(np.abs(out2), np.abs(t)) < 1e-15


def test_constraint_regions_first_claim_and_fixed_c(oracle):
    n = 16
    rng = np.random.default_rng(43)
    a = np.zeros((n, n)); a[:, :10] = 1
    b = np.zeros((n, n)); b[:, 6:] = 1           # overlaps a in columns 6..9
    ka, kb = 0.5, 3.0
    phi = -rng.uniform(0.05, 1.0, (n, n))
    t = np.where(a > 0, np.exp((ka + 1j) * phi), np.exp((kb + 1j) * phi))
    tc = oracle.OracleConstraint(n, n)
    i = tc.add_complex_constraint(a, 0.5, 0.5)
    j = tc.add_complex_constraint(b, 0.5, 0.5)
    out = tc.apply(t.copy())
    # region b holds columns 10.. only, so both regions are homogeneous
    assert abs(tc.c_mean(i) - ka) < 1e-13 and abs(tc.c_mean(j) - kb) < 1e-12
    assert rel(out, t) < 1e-14
    # a fixed c overrides the mean: alpha1 = 1 gives exp(c phi)
    tc2 = oracle.OracleConstraint(n, n)
    tc2.set_enforce_unity(False)
    tc2.add_complex_constraint(np.ones((n, n)), 1.0, 0.0, fixed_c=2.0)
    out2 = tc2.apply(t.copy())
    assert tc2.c_mean(0) == 2.0
    assert rel(np.abs(out2), np.exp(2.0 * phi)) < 1e-13


def test_fresnel_support_applies_constraint_to_transmission(oracle):
    """FresnelCDI::apply_support with a constraint: support, then the
    transmission function T = (esw + Ls)/Ls is constrained and written back
    as esw = (T - 1) Ls.  A homogeneous T inside a full support is unchanged."""
    n = 32
    o, x, wf, rng = fresnel(oracle, n)
    Ls = o.get_illumination_at_sample()
    T = homogeneous(n, 0.8, rng)
    esw = (T - 1) * Ls
    o.set_support(np.ones((n, n)))
    tc = oracle.OracleConstraint(n, n)
    tc.add_complex_constraint(np.ones((n, n)), 0.6, 0.6)
    o.set_complex_constraint(tc)
    e2 = esw.copy()
    o.apply_support(e2)
    assert rel(e2, esw) < 1e-13
    # a non-homogeneous T moves, and get_transmission_function applies the constraint
    T2 = T * np.exp(0.3j * rng.uniform(size=(n, n)))
    e3 = (T2 - 1) * Ls
    got = o.get_transmission_function(e3)
    assert rel(got, T2) > 1e-3
    o.apply_support(e3)
    tc0 = oracle.OracleConstraint(n, n)  # the same constraint applied by hand
    tc0.add_complex_constraint(np.ones((n, n)), 0.6, 0.6)
    assert rel(e3, (tc0.apply(T2.copy()) - 1) * Ls) < 1e-12


# ------------------------------------------------------- PhaseDiverseCDI ----
# Pinned by structure: the global object is the bounding box of the frames;
# a single frame with unit weights (series, beta 1) makes the big iteration
# the frame's own iteration; two identical frames in parallel mode give
# (1 - beta) object + beta frame; weights are alpha (|L_s|/max)^gamma on the
# support; 2x sub-pixel sampling maps constants to constants.

def planar_frame(oracle, n=32, seed=0, support=14):
    rng = np.random.default_rng(seed)
    obj = np.zeros((n, n), complex)
    s0 = n // 2 - support // 2
    obj[s0 + 2:s0 + support - 2, s0 + 2:s0 + support - 2] = rng.uniform(0.5, 1, (support - 4, support - 4))
    I = np.abs(np.fft.fftshift(np.fft.fft2(obj)) / n) ** 2
    sup = np.zeros((n, n)); sup[s0:s0 + support, s0:s0 + support] = 1
    x = (sup * np.exp(1j * rng.uniform(0, 2 * np.pi, (n, n)))).astype(complex)
    o = oracle.OracleCDI(x, oracle.PLANAR)
    o.set_support(sup)
    o.set_intensity(I)
    o.set_algorithm(oracle.HIO)
    return o, x, sup, I


def test_phase_diverse_geometry(oracle):
    pd = oracle.OraclePhaseDiverse()
    a, *_ = planar_frame(oracle, 32, 1)
    b, *_ = planar_frame(oracle, 32, 2)
    pd.add_new_position(a, 0, 0)
    pd.add_new_position(b, -8, 5)
    # frame at (x, y) covers object pixels local - (x, y) - origin: the box
    # of [-x, -x + 32) over both frames is [0, 40) x [-5, 32)
    assert pd.size() == (40, 37) and pd.origin() == (0, -5)
    pd2 = oracle.OraclePhaseDiverse()
    pd2.add_new_position(a, 0, 0)
    pd2.add_new_position(b, 10.5, -4.25)
    nx, ny = pd2.size()
    xm, ym = pd2.origin()
    for f, (x, y) in ((a, (0, 0)), (b, (10.5, -4.25))):
        for l in (1, 30):   # the interior pixels land inside the object
            assert 0 <= int(l - x - xm) < nx and 0 <= int(l - y - ym) < ny


def test_phase_diverse_single_frame_is_the_frame_iteration(oracle):
    o, x, sup, I = planar_frame(oracle, 32, 3)
    ref_x = x.copy()
    ref = oracle.OracleCDI(ref_x, oracle.PLANAR)
    ref.set_support(sup)
    ref.set_intensity(I)
    ref.set_algorithm(oracle.HIO)
    pd = oracle.OraclePhaseDiverse(beta=1.0)
    pd.add_new_position(o, 0, 0)
    assert np.array_equal(pd.get_weight(0), sup)
    pd.initialise_estimate()
    t = pd.get_transmission()
    inner = sup.astype(bool)
    # obj + w f - w obj with w = 1: f up to the rounding of (f + obj) - obj
    assert np.abs(t[inner] - x[inner]).max() < 1e-15 and np.all(t[~inner] == 1)
    for _ in range(3):
        pd.iterate()
        ref.iterate()
    assert rel(x, ref_x) < 1e-13
    assert np.abs(pd.get_transmission()[inner] - ref_x[inner]).max() < 1e-13


def test_phase_diverse_parallel_mix(oracle):
    beta = 0.7
    o1, x1, sup, I = planar_frame(oracle, 32, 4)
    o2, x2, _, _ = planar_frame(oracle, 32, 4)
    pd = oracle.OraclePhaseDiverse(beta=beta, parallel=True)
    pd.add_new_position(o1, 0, 0)
    pd.add_new_position(o2, 0, 0)
    assert np.allclose(pd.get_weight(0), sup * beta / 2, rtol=1e-15, atol=0)
    pd.initialise_estimate()
    before = pd.get_transmission()
    pd.iterate()
    # each frame ran one HIO step from the object; both are identical
    assert np.array_equal(x1, x2)
    inner = sup.astype(bool)
    expect = (1 - beta) * before + beta * x1
    after = pd.get_transmission()
    assert np.abs(after[inner] - expect[inner]).max() < 1e-14
    assert np.all(after[~inner] == 1)


def test_phase_diverse_fresnel_weights(oracle):
    n = 32
    o, x, wf, rng = fresnel(oracle, n)
    sup = np.zeros((n, n)); sup[6:26, 8:24] = 1
    o.set_support(sup)
    pd = oracle.OraclePhaseDiverse(beta=0.5, gamma=1.7)
    pd.add_new_position(o, 0, 0, alpha=2.0)
    m = np.abs(o.get_illumination_at_sample())
    expect = np.where(sup > 0, 2.0 * (m / m.max()) ** 1.7 * 0.5, 0.0)
    assert np.abs(pd.get_weight(0) - expect).max() < 1e-14


def test_phase_diverse_subpixel_constant(oracle):
    o, x, sup, I = planar_frame(oracle, 16, 5, support=8)
    x[:] = (0.25 - 0.5j) * sup
    pd = oracle.OraclePhaseDiverse(beta=1.0, granularity=2)
    pd.add_new_position(o, 0, 0)
    assert pd.size() == (32, 32)
    pd.initialise_estimate()
    t = pd.get_transmission()
    # interior support pixels (whose 2x2 neighbourhood is inside the support)
    # give their constant to every sub-pixel point of their block
    s0 = 4
    for i_ in range(s0 + 1, s0 + 7):
        for j_ in range(s0 + 1, s0 + 7):
            blk = t[2 * i_:2 * i_ + 2, 2 * j_:2 * j_ + 2]
            assert np.abs(blk - (0.25 - 0.5j)).max() < 1e-15
    # update_from_object averages each block back: a big iteration starts
    # the frame from the constant
    pd2 = oracle.OraclePhaseDiverse(beta=1.0, granularity=2)
    o3, x3, sup3, _ = planar_frame(oracle, 16, 6, support=8)
    pd2.add_new_position(o3, 0, 0)
    pd2.initialise_estimate()
    tt = pd2.get_transmission()
    tt[:] = 0.5 + 0.25j
    assert pd2.set_transmission(tt) == 0
    pd2.set_iterations_per_cycle(0)
    pd2.iterate()
    assert np.abs(x3[sup3 > 0] - (0.5 + 0.25j)).max() < 1e-15


def test_intensity_autocorrelation_is_object_autocorrelation(oracle):
    """PlanarCDI::get_intensity_autocorrelation = |IDFT(I)| centred and scaled
    (the (1+i)/sqrt(2) factor has unit modulus): by Wiener-Khinchin the
    centred cyclic autocorrelation of the object, computed here by brute force."""
    n = 16
    rng = np.random.default_rng(44)
    obj = np.zeros((n, n))
    obj[5:11, 4:9] = rng.uniform(0.2, 1.0, (6, 5))
    I = np.abs(np.fft.fftshift(np.fft.fft2(obj)) / n) ** 2
    o = oracle.OracleCDI(np.zeros((n, n), complex), oracle.PLANAR)
    o.set_intensity(I)
    got = o.get_intensity_autocorrelation()
    R = np.zeros((n, n))
    for kx in range(n):
        for ky in range(n):
            R[kx, ky] = np.sum(np.roll(np.roll(obj, -kx, 0), -ky, 1) * obj)
    # I = |DFT obj|^2 / N, so the unnormalised IDFT of I is R; invert(true)
    # centres it and divides by sqrt(N) = n
    expect = np.fft.fftshift(np.abs(R)) / n
    assert np.abs(got - expect).max() <= 1e-12 * expect.max()


def test_set_transmission_function_stays_in_the_field(oracle):
    """FresnelCDI::set_transmission_function loops j < nx
    (NADIA/src/FresnelCDI.c++:270-283).  For nx < ny the columns j >= nx
    keep their values; for nx > ny the reference's loop would run past the
    array: the oracle writes every pixel once and nothing beyond the field
    (guard elements after the estimate stay untouched)."""
    rng = np.random.default_rng(23)
    for nx, ny in ((32, 64), (64, 32)):
        wf = rng.random((nx, ny)) + 1j * rng.random((nx, ny)) + 0.5
        buf = np.full(nx * ny + 512, 7.0 + 7.0j)
        x = buf[:nx * ny].reshape(nx, ny)
        x[:] = 0
        o = oracle.OracleCDI(x, oracle.FRESNEL)
        o.fresnel_init(wf, 1.55e-10, 0.909, 1.6e-3, 13.5e-6, 1.0)
        o.set_support(np.ones((nx, ny)))
        o.set_intensity(np.abs(wf) ** 2)
        o.initialise_estimate(0)
        assert np.all(buf[nx * ny:] == 7.0 + 7.0j), (nx, ny)
        if nx < ny:
            assert np.all(x[:, nx:] == 0), (nx, ny)  # never written
        else:
            assert np.count_nonzero(x) > 0.9 * x.size
