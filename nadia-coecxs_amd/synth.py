"""Synthetic inputs for the reconstruction workloads (host side, numpy).

Every random draw uses numpy's PCG64 generator (numpy.random.default_rng)
with the seed named in each function; Poisson sampling is numpy's
Generator.poisson.  The far-field detector model is NADIA's centred unitary
propagator, F x = fftshift(DFT x)/sqrt(N) (NADIA/src/PlanarCDI.c++:65-68);
the Fresnel models follow NADIA/src/FresnelCDI.c++:57-126 and
NADIA/src/FresnelCDI_WF.c++:17-175.  Choices the workload descriptions leave
open are fixed here and named in the docstrings.
"""
import numpy as np


def _gauss_smooth(a, sigma):
    """Separable Gaussian smoothing, kernel truncated at 4 sigma, zero boundary."""
    r = int(np.ceil(4 * sigma))
    x = np.arange(-r, r + 1)
    k = np.exp(-x ** 2 / (2.0 * sigma ** 2))
    k /= k.sum()
    out = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), 0, a)
    out = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), 1, out)
    return out


def far_field(obj):
    """|F obj|^2 with F = fftshift(DFT)/sqrt(N)."""
    n = obj.size
    return np.abs(np.fft.fftshift(np.fft.fft2(obj)) / np.sqrt(n)) ** 2


def planar_block(n=256, block=96, support=110, peak=1e6, sigma=2.0, seed_amp=0, seed_poisson=0,
                 seed_phase=1):
    """Centred block object (the 256^2 HIO/ER workload at n=256).

    Object: centred block x block, amplitude U[0.5, 1] (seed_amp), Gaussian
    smoothing sigma px (kernel truncated at 4 sigma, zero boundary), zero
    phase, zero padded to n x n.  Intensity |F obj|^2 scaled to `peak` and
    Poisson sampled (seed_poisson).  Support: support x support centred square
    [n/2 - s/2, n/2 + s/2).  Initial guess: support * exp(i theta), theta
    U[0, 2 pi) (seed_phase).
    Returns dict(intensity, support, x0, obj).
    """
    rng = np.random.default_rng(seed_amp)
    obj = np.zeros((n, n))
    b0 = n // 2 - block // 2
    obj[b0:b0 + block, b0:b0 + block] = rng.uniform(0.5, 1.0, (block, block))
    obj = _gauss_smooth(obj, sigma)
    I = far_field(obj)
    I *= peak / I.max()
    I = np.random.default_rng(seed_poisson).poisson(I).astype(np.float64)
    sup = np.zeros((n, n))
    s0 = n // 2 - support // 2
    sup[s0:s0 + support, s0:s0 + support] = 1.0
    theta = np.random.default_rng(seed_phase).uniform(0, 2 * np.pi, (n, n))
    x0 = sup * np.exp(1j * theta)
    return dict(intensity=I, support=sup, x0=x0, obj=obj)


def planar_star(n=2048, radius=400, support=900, beamstop=16, peak=1e8, ngauss=20, seed=0):
    """Complex object in a star-shaped region (the 2048^2 workload at n=2048).

    Region: r(theta) <= R (1 + 0.15 sin(5 theta + a) + 0.05 sin(3 theta + b)),
    R = radius, a, b ~ U[0, 2 pi).  Amplitude U[0.3, 1] inside.  Phase: sum of
    ngauss Gaussians (sigma U[10, 60] px scaled by n/2048, centres inside the
    region, amplitudes U[-1, 1]) affinely scaled to [-pi, pi] over the region.
    Intensity Poisson at `peak`; central beamstop x beamstop pixels masked
    (beam-stop array value 0 there, 1 elsewhere).  Support: centred square of
    side `support`.  Initial guess: support * exp(i theta), theta U[0, 2 pi).
    Sizes scale with n/2048 when n differs.  Seeds: seed (shape, amplitude,
    phase), seed+1 (Poisson), seed+2 (initial phase).
    """
    rng = np.random.default_rng(seed)
    s = n / 2048.0
    R = radius * s
    yy, xx = np.mgrid[0:n, 0:n].astype(np.float64)
    cx = cy = (n - 1) / 2.0
    dx, dy = xx - cx, yy - cy
    th = np.arctan2(dy, dx)
    rr = np.hypot(dx, dy)
    a, b = rng.uniform(0, 2 * np.pi, 2)
    rlim = R * (1 + 0.15 * np.sin(5 * th + a) + 0.05 * np.sin(3 * th + b))
    region = rr <= rlim
    amp = np.where(region, rng.uniform(0.3, 1.0, (n, n)), 0.0)
    phase = np.zeros((n, n))
    inside = np.argwhere(region)
    for _ in range(ngauss):
        c = inside[rng.integers(len(inside))]
        sg = rng.uniform(10, 60) * s
        A = rng.uniform(-1, 1)
        phase += A * np.exp(-((yy - c[0]) ** 2 + (xx - c[1]) ** 2) / (2 * sg * sg))
    pr = phase[region]
    phase = (phase - pr.min()) / max(pr.max() - pr.min(), 1e-30) * 2 * np.pi - np.pi
    obj = amp * np.exp(1j * phase) * region
    I = far_field(obj)
    I *= peak / I.max()
    I = np.random.default_rng(seed + 1).poisson(I).astype(np.float64)
    beam = np.ones((n, n))
    b0 = n // 2 - beamstop // 2
    beam[b0:b0 + beamstop, b0:b0 + beamstop] = 0.0
    sup = np.zeros((n, n))
    sside = int(round(support * s))
    s0 = n // 2 - sside // 2
    sup[s0:s0 + sside, s0:s0 + sside] = 1.0
    theta = np.random.default_rng(seed + 2).uniform(0, 2 * np.pi, (n, n))
    x0 = sup * np.exp(1j * theta)
    return dict(intensity=I, support=sup, x0=x0, beam_stop=beam, obj=obj)


def batched_jobs(count=64, sizes=(256, 384, 512, 768, 1024), seed=300):
    """Independent block-object jobs of uneven sizes (the 64-job batch at the
    defaults): size n drawn uniformly from `sizes` (seed), then each job as
    planar_block with block side round(0.375 n) (96 px at 256), support side
    round(0.4 n), seeds seed+1+j (amplitude and Poisson) and seed+10001+j
    (initial phase).  Returns a list of dicts with an extra key n."""
    ns = np.random.default_rng(seed).choice(np.asarray(sizes), size=count)
    jobs = []
    for j, n in enumerate(ns):
        n = int(n)
        d = planar_block(n=n, block=int(round(0.375 * n)), support=int(round(0.4 * n)), seed_amp=seed + 1 + j,
                         seed_poisson=seed + 1 + j, seed_phase=seed + 10001 + j)
        d["n"] = n
        jobs.append(d)
    return jobs


def shrinkwrap_sigma(call, ncalls=50, start=3.0, end=1.5):
    """Gaussian width of shrinkwrap call `call` (0-based): linear start->end over ncalls."""
    if ncalls <= 1:
        return end
    t = min(call, ncalls - 1) / (ncalls - 1)
    return start + (end - start) * t


# ---------------------------------------------------------------- Fresnel ---

def _invert(a):
    return np.fft.fftshift(a) / np.sqrt(a.size)


def fresnel_phase(nx, ny, wavelength, zfd, zfs, pixel):
    """Quadratic phase of FresnelCDI (NADIA/src/FresnelCDI.c++:67-85)."""
    xm, ym = (nx - 1) / 2.0, (ny - 1) / 2.0
    zsd = zfd - zfs
    factor = pixel * pixel * np.pi / wavelength * ((1 / zfd) - (1 / zsd))
    i = np.arange(nx)[:, None]
    j = np.arange(ny)[None, :]
    return factor * ((xm - i) ** 2 + (ym - j) ** 2)


def fresnel_to_detector(c, phi):
    return np.fft.fft2(_invert(c)) * np.exp(-1j * phi)


def fresnel_from_detector(c, phi):
    return _invert(np.fft.ifft2(c * np.exp(1j * phi)) * c.size)


def wf_coefficient(nx, ny, wavelength, z12, z23, pixel):
    """(sin phi - i cos phi)/sqrt(N) of FresnelCDI_WF (NADIA/src/FresnelCDI_WF.c++:36-58)."""
    xm, ym = (nx - 1) / 2.0, (ny - 1) / 2.0
    sx = wavelength * z23 / (pixel * nx)
    sy = wavelength * z23 / (pixel * ny)
    i = np.arange(nx)[:, None]
    j = np.arange(ny)[None, :]
    rho2 = (sx * (xm - i)) ** 2 + (sy * (ym - j)) ** 2
    phi = (rho2 / z12 + rho2 / z23) * (np.pi / wavelength)
    return (np.sin(phi) - 1j * np.cos(phi)) / np.sqrt(nx * ny)


def wf_to_detector(c, coef):
    t = np.fft.fft2(c) * np.conj(coef)
    return np.fft.fft2(_invert(t))


def fresnel_config(n=1024, wavelength=1.55e-10, zfs=1.6e-3, zfd=0.909, pixel=13.5e-6, z12=16.353e-3,
                   pupil_diameter=160e-6, ripple=0.2, ndisks=12, disk_radius=20, mean_photons=1e7,
                   support=500, seed=0):
    """Fresnel workload (1024^2 at the defaults).

    Illumination: a pupil disk of diameter pupil_diameter in the zone-plate
    plane (pixel (z12/z23) * pixel, centred at the integer (n-1)/2 like
    FresnelCDI_WF::set_support) with i.i.d. phase ripple U[-ripple, ripple]
    (uncorrelated per pixel), taken to the detector by FresnelCDI_WF's
    propagate_to_detector (z23 = zfd) -> white field L_d.  White-field
    intensity |L_d|^2 scaled to `mean_photons` mean.  Object: transmission
    exp(-mu t + i phi) with t the union of `ndisks` disks of radius
    disk_radius (40 px diameter) in the central support, mu t U[0, 0.3] and
    phi U[0, 1.2] per disk.  ESW = (T - 1) L_s, L_s = from_detector(L_d);
    intensity |to_detector(ESW) + L_d|^2 scaled to `mean_photons` mean and
    Poisson sampled.  Support: centred square of side `support` (scaled by
    n/1024).  Seeds: seed (pupil ripple, disks), seed+1 (Poisson).
    """
    rng = np.random.default_rng(seed)
    s = n / 1024.0
    z23 = zfd
    dz = (z12 / z23) * pixel
    c = (n - 1) // 2
    i = np.arange(n)[:, None] - c
    j = np.arange(n)[None, :] - c
    r = dz * np.sqrt(i * i + j * j)
    pupil = (r < pupil_diameter / 2.0) * np.exp(1j * rng.uniform(-ripple, ripple, (n, n)))
    coef = wf_coefficient(n, n, wavelength, z12, z23, pixel)
    Ld = wf_to_detector(pupil, coef)
    wf_int = np.abs(Ld) ** 2
    scale = mean_photons / wf_int.mean()
    Ld = Ld * np.sqrt(scale)
    wf_int = np.abs(Ld) ** 2
    phi = fresnel_phase(n, n, wavelength, zfd, zfs, pixel)
    Ls = fresnel_from_detector(Ld, phi)
    T = np.ones((n, n), complex)
    side = int(round(support * s))
    lo = n // 2 - side // 2 + int(disk_radius * s) + 2
    hi = n // 2 + side // 2 - int(disk_radius * s) - 2
    yy, xx = np.mgrid[0:n, 0:n]
    for _ in range(ndisks):
        cy, cx = rng.integers(lo, hi, 2)
        mt = rng.uniform(0, 0.3)
        ph = rng.uniform(0, 1.2)
        m = (yy - cy) ** 2 + (xx - cx) ** 2 <= (disk_radius * s) ** 2
        T[m] = np.exp(-mt + 1j * ph)
    esw = (T - 1) * Ls
    I = np.abs(fresnel_to_detector(esw, phi) + Ld) ** 2
    I *= mean_photons / I.mean()
    I = np.random.default_rng(seed + 1).poisson(I).astype(np.float64)
    sup = np.zeros((n, n))
    s0 = n // 2 - side // 2
    sup[s0:s0 + side, s0:s0 + side] = 1.0
    return dict(intensity=I, support=sup, white_field=Ld, white_field_intensity=wf_int, pupil=pupil,
                transmission=T, esw=esw, wavelength=wavelength, zfd=zfd, zfs=zfs, pixel=pixel, z12=z12,
                z23=z23, pupil_diameter=pupil_diameter)


def disk_field(n, count, rng, radius=(2.0, 5.0)):
    """n x n binary field: 1 inside any of `count` disks (radius uniform in
    `radius` px, uniform real centres), 0 elsewhere -- isolated features of one
    value on an exactly zero background."""
    yy, xx = np.mgrid[0:n, 0:n]
    g = np.zeros((n, n))
    for _ in range(count):
        cx, cy = rng.uniform(0, n, 2)
        r = rng.uniform(*radius)
        g[(xx - cx) ** 2 + (yy - cy) ** 2 < r * r] = 1.0
    return g


def ptycho_frames(m=64, crops=((0, 0), (14, 9), (27, 20)), seed=7):
    """Overlapping planar frames cut from one complex object, for position
    refinement.

    Object: (m + max crop) square, magnitude 1.05 - 0.65 g with g a field of
    disks (disk_field: absorbing features of one value), phase 0.8 g.  The
    cross-correlation alignment looks at 1 - |T| where |T| < 1 and weights
    exactly those pixels: a background magnitude clear of 1 keeps that set
    independent of rounding and leaves only the features, on which the
    correlation ratio peaks at the true offset.  Frame k is the m x m crop at crops[k] = (X, Y), i.e. frame
    pixel (i, j) is object pixel (i + X, j + Y); in PhaseDiverseCDI
    coordinates its position is (-X, -Y).  Each frame comes with the exact
    noise-free intensity |F frame|^2 (unitary centred transform) and a full
    support, so the frame is a fixed point of its own ER / HIO iteration.
    Returns dict(object, frames, intensities, supports, positions).
    """
    rng = np.random.default_rng(seed)
    big = m + max(max(c) for c in crops)
    g = disk_field(big, big // 4, rng)
    obj = (1.05 - 0.65 * g) * np.exp(0.8j * g)
    frames, ints = [], []
    for X, Y in crops:
        f = obj[X:X + m, Y:Y + m].copy()
        frames.append(f)
        ints.append(far_field(f))
    return dict(object=obj, frames=frames, intensities=ints, supports=[np.ones((m, m)) for _ in crops],
                positions=[(-float(X), -float(Y)) for X, Y in crops])


def reference_planar_example(npz_path, threshold=0.04):
    """The reference's example input for PlanarCDI (the 1024 x 1024 uint16
    diffraction pattern NADIA/examples/image_files/planar_data.tif, kept as
    tests/golden/planar_data.npz), oriented as NADIA's TIFF reader loads it
    (x = column, y = row: intensity[x, y] = pixel at row y, column x).  The
    example's support file is not part of the reference's distribution; the
    support here is the set where the intensity autocorrelation
    |IDFT(I)| (centred) exceeds `threshold` of its maximum -- about twice the
    object's extent, which the example's shrinkwrap steps then tighten.
    Returns dict(intensity, support)."""
    img = np.load(npz_path)["rows_cols"].astype(np.float64)
    inten = np.ascontiguousarray(img.T)
    a = np.abs(np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(inten))))
    support = (a > threshold * a.max()).astype(np.float64)
    return dict(intensity=inten, support=support)


# NADIA/examples/PlanarCDI_example.c:40-222: per cycle 50 ER (shrinkwrap
# sigma 1.5 every 50), relaxation 0.5 then 100 HIO and 50 ER (shrinkwrap
# sigma 2.0 every 50), threshold 0.1, two cycles, 4 best estimates kept
REFERENCE_EXAMPLE_SCHEDULE = dict(cycles=2, er1=50, hio=100, er2=50, shrinkwrap_every=50, sigma_er1=1.5,
                                  sigma_rest=2.0, threshold=0.1, beta=0.5, n_best=4, seed=0)


def run_reference_example(obj, alg, iterate, shrinkwrap, set_beta, cycles=None, sched=REFERENCE_EXAMPLE_SCHEDULE):
    """Drives `obj` (anything with the callables given) through the example's
    loop: iterate(), shrinkwrap(sigma, threshold), set_beta(b), alg(name).
    Returns the errors after each iteration."""
    s = sched
    errs = []
    cyc = s["cycles"] if cycles is None else cycles
    for _ in range(cyc):
        alg("ER")
        for i in range(s["er1"]):
            errs.append(iterate())
            if i % s["shrinkwrap_every"] == s["shrinkwrap_every"] - 1:
                shrinkwrap(s["sigma_er1"], s["threshold"])
        set_beta(s["beta"])
        alg("HIO")
        for i in range(s["er1"], s["er1"] + s["hio"]):
            errs.append(iterate())
            if i % s["shrinkwrap_every"] == s["shrinkwrap_every"] - 1:
                shrinkwrap(s["sigma_rest"], s["threshold"])
        alg("ER")
        for i in range(s["er1"] + s["hio"], s["er1"] + s["hio"] + s["er2"]):
            errs.append(iterate())
            if i % s["shrinkwrap_every"] == s["shrinkwrap_every"] - 1:
                shrinkwrap(s["sigma_rest"], s["threshold"])
    return errs
