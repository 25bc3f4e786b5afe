"""GPU parity of TransmissionConstraint (nadia_constraint) against the CPU
oracle: the constraint alone on a field, and inside apply_support during
PlanarCDI and FresnelCDI iterations: ER on the fused passes with the
constraint between iterations, other algorithms and the plain path
unfused."""
import numpy as np
import pytest

from test_gpu_fresnel import fresnel_pair

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def field(n, seed):
    rng = np.random.default_rng(seed)
    mag = rng.uniform(0.2, 1.4, (n, n))
    return mag * np.exp(1j * rng.uniform(-np.pi, np.pi, (n, n)))


def regions(n):
    a = np.zeros((n, n))
    a[n // 8:n // 2, n // 8:5 * n // 8] = 1
    b = np.zeros((n, n))
    b[3 * n // 8:7 * n // 8, n // 4:3 * n // 4] = 1  # overlaps a: a keeps the shared pixels
    return a, b


def make_pair(nd, oracle, n, double, flip=(True, 1), unity=True, fixed=None, regs=None):
    g = nd.TransmissionConstraint(n, n)
    o = oracle.OracleConstraint(n, n, double=double)
    a, b = regs if regs is not None else regions(n)
    for tc in (g, o):
        tc.set_charge_flipping(*flip)
        tc.set_enforce_unity(unity)
        tc.add_complex_constraint(a, 0.6, 0.3)
        tc.add_complex_constraint(b, 0.2, 0.9, fixed_c=fixed)
    return g, o


@pytest.mark.parametrize("double,tol", [(True, 1e-12), (False, 1e-5)], ids=["fp64", "fp32"])
@pytest.mark.parametrize("flip,unity,fixed", [((True, 1), True, None), ((True, -1), False, 0.7),
                                              ((False, 1), True, None)])
def test_constraint_matches_oracle(nd, oracle, double, tol, flip, unity, fixed):
    n = 96
    cdt = np.complex128 if double else np.complex64
    t = field(n, 3).astype(cdt)
    g, o = make_pair(nd, oracle, n, double, flip, unity, fixed)
    f = nd.Complex2D(n, n, double=double, data=t)
    g.apply_constraint(f)
    to = t.copy()
    o.apply(to)
    assert rel(f.get(), to) <= tol
    for k in (0, 1):
        assert abs(g.c_mean(k) - o.c_mean(k)) <= tol * abs(o.c_mean(k))


def test_custom_constraint_runs_after_builtin(nd, oracle):
    n = 64
    t = field(n, 4)
    g, o = make_pair(nd, oracle, n, True)
    g.set_custom_constraint(lambda a: np.multiply(a, 0.5, out=a))
    f = nd.Complex2D(n, n, data=t)
    g.apply_constraint(f)
    to = t.copy()
    o.apply(to)
    assert rel(f.get(), 0.5 * to) <= 1e-12


@pytest.mark.parametrize("path", [0, 1], ids=["fused", "plain"])
@pytest.mark.parametrize("alg", ["ER", "HIO"])
def test_planar_iterations_with_constraint(nd, oracle, synth, alg, path):
    n = 128
    d = synth.planar_block(n=n, block=48, support=56)
    est = nd.Complex2D(n, n, data=d["x0"])
    g = nd.PlanarCDI(est)
    g.set_path(path)
    xo = d["x0"].astype(np.complex128).copy()
    o = oracle.OracleCDI(xo, oracle.PLANAR)
    # two overlapping regions inside the support (log|x| is finite there)
    s0 = n // 2 - 28
    a = np.zeros((n, n))
    a[s0 + 2:s0 + 30, s0 + 2:s0 + 54] = 1
    b = np.zeros((n, n))
    b[s0 + 20:s0 + 54, s0 + 2:s0 + 54] = 1
    tg, to = make_pair(nd, oracle, n, True, regs=(a, b))
    g.set_complex_constraint(tg)
    o.set_complex_constraint(to)
    for obj in (g, o):
        obj.set_support(d["support"])
        obj.set_intensity(d["intensity"])
        obj.set_algorithm(getattr(oracle, alg))
    eg = g.iterate(12)
    eo = o.iterate_n(12)
    assert (np.abs(eg - eo) / eo).max() <= 1e-10
    assert rel(est.get(), xo) <= 1e-10
    assert abs(tg.c_mean(0) - to.c_mean(0)) <= 1e-10 * abs(to.c_mean(0))


def fresnel_constrained(nd, oracle, synth, double=True, path=0):
    """A Fresnel pair with one complex-constraint region over the support,
    started from a transmission function with phases well away from zero:
    initialise_estimate's real T makes the region constant c = sum|log|T|| /
    sum|arg T| divide by rounding-level phases (ill-conditioned on any two
    implementations)."""
    d, est, g, xo, o = fresnel_pair(nd, oracle, synth, double=double, path=path)
    n = 128
    tg = nd.TransmissionConstraint(n, n)
    tcs = oracle.OracleConstraint(n, n)
    for tc in (tg, tcs):
        tc.add_complex_constraint(d["support"], 0.5, 0.5)
    g.set_complex_constraint(tg)
    o.set_complex_constraint(tcs)
    rng = np.random.default_rng(5)
    T0 = np.exp((0.3 + 1j) * -rng.uniform(0.1, 0.4, (n, n)))
    g.set_transmission_function(T0)
    o.set_transmission_function(T0)
    return est, g, xo, o, tg, tcs


@pytest.mark.parametrize("path", [0, 1], ids=["fused", "plain"])
def test_fresnel_iterations_with_constraint(nd, oracle, synth, path):
    d, est, g, xo, o = fresnel_pair(nd, oracle, synth, path=path)
    n = 128
    tg = nd.TransmissionConstraint(n, n)
    tcs = oracle.OracleConstraint(n, n)
    inside = d["support"]
    for tc in (tg, tcs):
        tc.add_complex_constraint(inside, 0.5, 0.5)
    g.set_complex_constraint(tg)
    o.set_complex_constraint(tcs)
    # start from a transmission function with phases well away from zero:
    # initialise_estimate's real T makes the region constant c = sum|log|T|| /
    # sum|arg T| divide by rounding-level phases (ill-conditioned on any two
    # implementations)
    rng = np.random.default_rng(5)
    phi = -rng.uniform(0.1, 0.4, (n, n))
    T0 = np.exp((0.3 + 1j) * phi)
    g.set_transmission_function(T0)
    o.set_transmission_function(T0)
    assert rel(est.get(), xo) <= 1e-13
    eg = g.iterate(8)
    eo = o.iterate_n(8)
    assert (np.abs(eg - eo) / eo).max() <= 1e-10
    assert rel(est.get(), xo) <= 1e-10
    # get_transmission_function applies the constraint, as in the reference
    assert rel(g.get_transmission_function(), o.get_transmission_function()) <= 1e-10


def test_fresnel_constraint_fp32(nd, oracle, synth):
    """FP32 (fused ER with the constraint) within the Fresnel FP32 bound
    (1e-4) of the FP64 oracle."""
    est, g, xo, o, tg, tcs = fresnel_constrained(nd, oracle, synth, double=False)
    eg = g.iterate(4)
    eo = o.iterate_n(4)
    assert (np.abs(eg - eo) / eo).max() <= 1e-4
    assert rel(est.get().astype(np.complex128), xo) <= 1e-4


def test_constrained_job_in_batch_matches_alone(nd, oracle, synth):
    """A constrained Fresnel ER job batched with an unconstrained planar job:
    both bit-identical to running alone."""
    est_a, g_a, _, _, tg_a, _ = fresnel_constrained(nd, oracle, synth)
    est_b, g_b, _, _, tg_b, _ = fresnel_constrained(nd, oracle, synth)
    e_a = g_a.iterate(5)
    d = synth.planar_block(n=64, block=24, support=30)
    est_p = nd.Complex2D(64, 64, data=d["x0"])
    p = nd.PlanarCDI(est_p)
    p.set_support(d["support"])
    p.set_intensity(d["intensity"])
    p.set_algorithm(nd.HIO)
    est_q = nd.Complex2D(64, 64, data=d["x0"])
    q = nd.PlanarCDI(est_q)
    q.set_support(d["support"])
    q.set_intensity(d["intensity"])
    q.set_algorithm(nd.HIO)
    e_q = q.iterate(5)
    b = nd.Batch([g_b, p])
    e = b.iterate(5)
    assert np.array_equal(e[:, 0], e_a) and np.array_equal(e[:, 1], e_q)
    assert np.array_equal(est_b.get(), est_a.get()) and np.array_equal(est_p.get(), est_q.get())


def test_fresnel_empty_constraint_round_trip(nd, oracle, synth):
    """A constraint with no region, no charge flipping and no unity cap still
    takes the exit wave through T and back (get_transmission_function then
    set_transmission_function), as the reference does."""
    d, est, g, xo, o = fresnel_pair(nd, oracle, synth)
    n = 128
    tg = nd.TransmissionConstraint(n, n)
    tcs = oracle.OracleConstraint(n, n)
    for tc in (tg, tcs):
        tc.set_charge_flipping(False)
        tc.set_enforce_unity(False)
    g.set_complex_constraint(tg)
    o.set_complex_constraint(tcs)
    g.initialise_estimate(3)
    o.initialise_estimate(3)
    eg = g.iterate(5)
    eo = o.iterate_n(5)
    assert (np.abs(eg - eo) / eo).max() <= 1e-10
    assert rel(est.get(), xo) <= 1e-10


def test_fresnel_custom_constraint_fused_matches_plain(nd, oracle, synth):
    """A host constraint function (it needs the T array) on the fused ER path
    agrees with the plain path."""
    outs = []
    for path in (0, 1):
        est, g, xo, o, tg, tcs = fresnel_constrained(nd, oracle, synth, path=path)
        tg.set_custom_constraint(lambda a: np.multiply(a, 0.999, out=a))
        e = g.iterate(4)
        outs.append((est.get(), e))
    assert rel(outs[0][0], outs[1][0]) <= 1e-10
    assert np.abs(outs[0][1] - outs[1][1]).max() <= 1e-10 * np.abs(outs[1][1]).max()


@pytest.mark.parametrize("nreg", [4, 5], ids=["4regions", "5regions"])
def test_planar_er_many_regions(nd, oracle, synth, nreg):
    """ER with 4 regions (summed inside the update pass) and with 5 (a
    separate reduction pass) against the oracle."""
    n = 128
    d = synth.planar_block(n=n, block=48, support=56)
    est = nd.Complex2D(n, n, data=d["x0"])
    g = nd.PlanarCDI(est)
    xo = d["x0"].astype(np.complex128).copy()
    o = oracle.OracleCDI(xo, oracle.PLANAR)
    s0 = n // 2 - 28
    tg = nd.TransmissionConstraint(n, n)
    to = oracle.OracleConstraint(n, n)
    for k in range(nreg):
        reg = np.zeros((n, n))
        reg[s0 + 2 + 10 * k:s0 + 12 + 10 * k, s0 + 2:s0 + 54] = 1
        for tc in (tg, to):
            tc.add_complex_constraint(reg, 0.3 + 0.1 * k, 0.6 - 0.1 * k)
    g.set_complex_constraint(tg)
    o.set_complex_constraint(to)
    for obj in (g, o):
        obj.set_support(d["support"])
        obj.set_intensity(d["intensity"])
        obj.set_algorithm(oracle.ER)
    eg = g.iterate(8)
    eo = o.iterate_n(8)
    assert (np.abs(eg - eo) / eo).max() <= 1e-10
    assert rel(est.get(), xo) <= 1e-10
    for k in range(nreg):
        assert abs(tg.c_mean(k) - to.c_mean(k)) <= 1e-10 * abs(to.c_mean(k))


@pytest.mark.parametrize("shape", [(128, 256), (256, 128)])
@pytest.mark.parametrize("path", [0, 1], ids=["fused", "plain"])
def test_rectangular_fresnel_with_constraint(nd, oracle, synth, shape, path):
    """Constrained FresnelCDI ER on rectangular fields: the region map, the
    transmission <-> exit-wave conversions (the reference's j < nx columns
    for nx < ny, every pixel for nx > ny) and the fused constraint passes,
    eight iterations within 1e-10 of the oracle."""
    nx, ny = shape
    d = synth.fresnel_config(n=256, pupil_diameter=160e-6 * 256 / 1024, disk_radius=20 * 256 / 1024 + 2,
                             ndisks=4, support=500)
    cut = lambda a: np.ascontiguousarray(a[:nx, :ny])
    est = nd.Complex2D(nx, ny, double=True)
    g = nd.FresnelCDI(est, cut(d["white_field"]).astype(np.complex128), d["wavelength"], d["zfd"], d["zfs"],
                      d["pixel"], 1.0)
    g.set_path(path)
    xo = np.zeros(shape, np.complex128)
    o = oracle.OracleCDI(xo, oracle.FRESNEL)
    o.fresnel_init(cut(d["white_field"]), d["wavelength"], d["zfd"], d["zfs"], d["pixel"], 1.0)
    for obj in (g, o):
        obj.set_support(cut(d["support"]))
        obj.set_intensity(cut(d["intensity"]))
    tg = nd.TransmissionConstraint(nx, ny)
    tcs = oracle.OracleConstraint(nx, ny)
    inside = cut(d["support"])
    for tc in (tg, tcs):
        tc.add_complex_constraint(inside, 0.5, 0.5)
    g.set_complex_constraint(tg)
    o.set_complex_constraint(tcs)
    rng = np.random.default_rng(6)
    T0 = np.exp((0.3 + 1j) * -rng.uniform(0.1, 0.4, shape))
    g.set_transmission_function(T0)
    o.set_transmission_function(T0)
    assert rel(est.get(), xo) <= 1e-13
    eg = g.iterate(8)
    eo = o.iterate_n(8)
    assert (np.abs(eg - eo) / eo).max() <= 1e-10
    assert rel(est.get(), xo) <= 1e-10
    assert rel(g.get_transmission_function(), o.get_transmission_function()) <= 1e-10
