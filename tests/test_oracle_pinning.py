"""Pinning of four oracle functions against computations written
independently of the oracle (numpy FFTs and closed forms):

* one FresnelCDI_WF iteration (NADIA/src/FresnelCDI_WF.c++:122-175) as the
  composition of numpy propagators with the closed-form coefficient and the
  support rule (NADIA/src/BaseCDI.c++:136-154), plus a constructed fixed
  point;
* a soft support (0 < s < 1) in ER and HIO (NADIA/src/BaseCDI.c++:136-154,
  297-387): x' = s P_F x, with P_F from numpy;
* which estimate update_n_best keeps (NADIA/src/BaseCDI.c++:390-413): the
  estimate after the iteration, filed under the error that iteration
  reported (the error of the estimate before it);
* PhaseDiverseCDI::iterate in series mode with two distinct frames at
  different positions (NADIA/src/PhaseDiverseCDI.c++:202-245, 721-1130),
  recomputed by hand: update_from_object, the frame's ER step from numpy,
  add_to_object with the series rule obj + w f - w obj, in frame order.

Each check is chosen so that a plausible mistake fails: the wrong
coefficient direction, a support multiplied where s >= 1, a dropped
fftshift, the pre-update estimate kept as the snapshot, frames merged in
parallel instead of in series.
"""
import numpy as np
import pytest


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


WF = dict(wavelength=1.55e-10, z12=16.353e-3, z23=0.909, pixel=13.5e-6)


def wf_coef(n):
    """(sin phi - i cos phi)/n with phi = (pi/lambda)(rho^2/z12 + rho^2/z23),
    rho measured from the array centre (n-1)/2 in the focal plane."""
    xm = (n - 1) / 2.0
    s = WF["wavelength"] * WF["z23"] / (WF["pixel"] * n)
    i = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    rho2 = (s * (xm - i)) ** 2 + (s * (xm - j)) ** 2
    phi = (rho2 / WF["z12"] + rho2 / WF["z23"]) * np.pi / WF["wavelength"]
    return (np.sin(phi) - 1j * np.cos(phi)) / n


def np_wf_from_detector(a, c):
    """backward DFT, centre shift / sqrt(N), x coefficient, backward DFT."""
    n = a.shape[0]
    u = np.fft.fftshift(np.fft.ifft2(a) * a.size) / n * c
    return np.fft.ifft2(u) * a.size


def np_wf_to_detector(b, c):
    """forward DFT, x conj(coefficient), centre shift / sqrt(N), forward DFT."""
    n = b.shape[0]
    u = np.fft.fft2(b) * np.conj(c)
    return np.fft.fft2(np.fft.fftshift(u) / n)


def np_support(x, s):
    """support_constraint: s == 0 -> 0, 0 < s < 1 -> scaled, s >= 1 -> kept."""
    out = x.copy()
    out[s == 0] = 0
    soft = (s > 0) & (s < 1)
    out[soft] = x[soft] * s[soft]
    return out


def np_modulus(y, q):
    """scale_intensity without a beam stop: |y| -> q keeping the phase,
    (q, 0) where |y| = 0; error sum (|y| - q)^2 / sum q^2."""
    m = np.abs(y)
    out = np.where(m > 0, y * (q / np.where(m > 0, m, 1.0)), q + 0j)
    return out, float(np.sum((m - q) ** 2) / np.sum(q * q))


def wf_object(oracle, n, diameter_px=12.0):
    x = np.zeros((n, n), complex)
    o = oracle.OracleCDI(x, oracle.FRESNEL_WF)
    o.wf_init(WF["wavelength"], WF["z12"], WF["z23"], WF["pixel"])
    dz = WF["z12"] / WF["z23"] * WF["pixel"]
    o.wf_set_support(diameter_px * dz, 1.0)
    return o, x


def test_wf_iteration_is_composition_of_numpy_propagators(oracle):
    n = 32
    o, x = wf_object(oracle, n)
    s = o.get_support()
    # the softened disk has all three support regimes
    assert (s == 0).any() and ((s > 0) & (s < 1)).any() and (s > 1).any()
    rng = np.random.default_rng(70)
    I = rng.uniform(0.5, 3.0, (n, n)) ** 2
    o.set_intensity(I)
    x[:] = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    x0 = x.copy()
    o.iterate()
    c = wf_coef(n)
    b = np_support(np_wf_from_detector(x0, c), s)
    y, err = np_modulus(np_wf_to_detector(b, c), np.sqrt(I))
    assert rel(x, y) < 1e-12
    assert abs(o.get_error() - err) < 1e-12 * err


def test_wf_iteration_fixed_point(oracle):
    """A zone-plate-plane field supported where s >= 1 (left unchanged by the
    support rule), taken to the detector, with I = |detector field|^2, is a
    fixed point of the white-field iteration with error 0."""
    n = 32
    o, x = wf_object(oracle, n)
    s = o.get_support()
    rng = np.random.default_rng(71)
    z = np.where(s >= 1, rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)), 0)
    d = np_wf_to_detector(z, wf_coef(n))
    o.set_intensity(np.abs(d) ** 2)
    x[:] = d
    o.iterate()
    assert rel(x, d) < 1e-12
    assert o.get_error() < 1e-24


def np_pf(x, q):
    """P_F for PlanarCDI: F x = fftshift(DFT x)/sqrt(N), modulus, F^-1."""
    n = np.sqrt(x.size)
    y, err = np_modulus(np.fft.fftshift(np.fft.fft2(x)) / n, q)
    return np.fft.ifft2(np.fft.ifftshift(y)) * n, err


@pytest.mark.parametrize("alg", ["ER", "HIO"])
def test_soft_support_in_iteration(oracle, alg):
    n = 32
    rng = np.random.default_rng(72)
    x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    x0 = x.copy()
    o = oracle.OracleCDI(x, oracle.PLANAR)
    I = rng.uniform(0.5, 4.0, (n, n)) ** 2
    o.set_intensity(I)
    raw = np.zeros((n, n))
    raw[6:26, 5:27] = rng.uniform(0.1, 2.0, (20, 22))
    raw[10, 10] = 2.0  # the maximum: s = raw / 2 has values in (0, 1) and 1
    o.set_support(raw)
    s = raw / raw.max()
    assert ((s > 0) & (s < 1)).sum() > 100 and (s == 1).any() and (s == 0).any()
    if alg == "HIO":
        o.set_algorithm(oracle.HIO)
    o.iterate()
    pf, err = np_pf(x0, np.sqrt(I))
    if alg == "ER":
        expect = np_support(pf, s)
    else:  # HIO, beta 0.9: (1 + b) Ps Pf x - Ps x - b Pf x + x
        expect = x0 - 0.9 * pf - np_support(x0, s) + 1.9 * np_support(pf, s)
    assert rel(x, expect) < 1e-13
    assert abs(o.get_error() - err) < 1e-13 * err


def test_n_best_keeps_post_update_estimates(oracle):
    """Snapshot k is the estimate after the iteration whose get_error() (the
    error of the estimate before that iteration) ranked k."""
    n = 32
    rng = np.random.default_rng(73)
    x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    xs = x.copy()
    I = rng.uniform(0.5, 4.0, (n, n)) ** 2
    sup = np.zeros((n, n))
    sup[8:24, 8:24] = 1
    steps, afters, errs = x.copy(), [], []
    o = oracle.OracleCDI(steps, oracle.PLANAR)
    on = oracle.OracleCDI(xs, oracle.PLANAR, n_best=3)
    for obj in (o, on):
        obj.set_intensity(I)
        obj.set_support(sup)
        obj.set_algorithm(oracle.HIO)
    for _ in range(10):
        o.iterate()
        afters.append(steps.copy())
        errs.append(o.get_error())
    on.iterate_n(10)
    assert np.array_equal(xs, steps)
    order = np.argsort(errs, kind="stable")
    assert len(set(errs)) == len(errs) and max(errs) < 1.0
    for k in range(3):
        snap, e = on.get_best_result(k)
        i = order[k]
        assert e == errs[i]
        assert np.array_equal(snap, afters[i])
        # and not the estimate the error was measured on
        before = afters[i - 1] if i > 0 else x
        assert rel(snap, before) > 1e-6


def test_phase_diverse_series_two_frames_by_hand(oracle):
    """Series mode, beta 0.6, two distinct ER frames at (0, 0) and (-5, 3):
    one big iteration recomputed frame by frame."""
    n, beta = 24, 0.6
    rng = np.random.default_rng(74)
    frames = []
    for f in range(2):
        x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        o = oracle.OracleCDI(x, oracle.PLANAR)
        I = rng.uniform(0.5, 4.0, (n, n)) ** 2
        sup = np.zeros((n, n))
        sup[4 + f:20 - f, 3:21] = 1
        o.set_intensity(I)
        o.set_support(sup)  # ER: the BaseCDI default
        frames.append((o, x, sup, np.sqrt(I)))
    pos = [(0.0, 0.0), (-5.0, 3.0)]
    pd = oracle.OraclePhaseDiverse(beta=beta, parallel=False)
    for (o, *_), (px, py) in zip(frames, pos):
        pd.add_new_position(o, px, py)
    nx, ny = pd.size()
    xm, ym = pd.origin()
    assert (nx, ny) == (29, 27) and (xm, ym) == (0, -3)
    ests = [x.copy() for _, x, _, _ in frames]
    pd.initialise_estimate()
    pd.iterate()

    def gidx(p):
        px, py = pos[p]
        gi = (np.arange(n) - px - xm).astype(int)
        gj = (np.arange(n) - py - ym).astype(int)
        return gi, gj

    w = [beta * sup for _, _, sup, _ in frames]  # series weights: beta on the support
    obj = np.ones((nx, ny), complex)

    def add(p, f):
        gi, gj = gidx(p)
        for i_ in range(1, n - 1):
            for j_ in range(1, n - 1):
                if w[p][i_, j_] != 0:
                    o_ = obj[gi[i_], gj[j_]]
                    obj[gi[i_], gj[j_]] = (w[p][i_, j_] * f[i_, j_] + o_) - w[p][i_, j_] * o_

    results = [e.copy() for e in ests]
    for p in range(2):
        add(p, results[p])
    for p in range(2):
        gi, gj = gidx(p)
        small = results[p].copy()
        m = w[p] != 0
        small[m] = obj[np.ix_(gi, gj)][m]
        pf, _ = np_pf(small, frames[p][3])
        results[p] = np_support(pf, frames[p][2])
        add(p, results[p])
    for p in range(2):
        assert rel(frames[p][1], results[p]) < 1e-13
    t = pd.get_transmission()
    assert np.abs(t - obj).max() < 1e-13
    # the second frame saw the first frame's update: merging in parallel
    # would give a different object
    assert np.abs(obj - 1).max() > 0.1
