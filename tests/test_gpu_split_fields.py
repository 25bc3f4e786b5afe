"""Split work fields: for PlanarCDI / FresnelCDI jobs with one work field and
no transmission constraint (ER at any size, other algorithms for rows of up
to 1024 points) the column pass leaves the projection in the second work
field, the update pass rebuilds x from the row transform kept in the first
one and stores x only when a sequence of iterations ends.  A job's
trajectory must not depend on how its iterations are split into calls or
batches, and must follow the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def job(nd, synth, n, seed, alg, double, n_best=0):
    d = synth.planar_block(n=n, block=3 * n // 8, support=7 * n // 16, seed_amp=seed, seed_poisson=seed,
                           seed_phase=seed + 1)
    cdt = np.complex128 if double else np.complex64
    est = nd.Complex2D(n, n, double=double, data=d["x0"].astype(cdt))
    g = nd.PlanarCDI(est, n_best=n_best)
    g.set_support(d["support"])
    g.set_intensity(d["intensity"])
    g.set_algorithm(alg)
    return d, est, g


@pytest.mark.parametrize("double", [True, False], ids=["fp64", "fp32"])
@pytest.mark.parametrize("alg", ["HIO", "ER", "RAAR"])
def test_trajectory_independent_of_call_split(nd, synth, double, alg):
    """7 iterations as one call, as 3 + 4, as seven single calls, and inside
    a batch with uneven counts per job (the job's own count split 2 + 5):
    bit-identical estimates and errors."""
    a = getattr(nd, alg)
    runs = []
    for split in ([7], [3, 4], [1] * 7):
        _, est, g = job(nd, synth, 512, 41, a, double)
        e = np.concatenate([g.iterate(k) for k in split])
        runs.append((est.get(), e))
    _, est, g = job(nd, synth, 512, 41, a, double)
    _, est2, g2 = job(nd, synth, 512, 42, a, double)
    _, est3, g3 = job(nd, synth, 512, 43, a, double)
    b = nd.Batch([g, g2, g3])
    b.iterate_each([2, 6, 1])
    b.iterate_each([5, 1, 3])
    runs.append((est.get(), None))
    x0, e0 = runs[0]
    for x, e in runs[1:]:
        assert np.array_equal(x, x0)
        if e is not None:
            assert np.array_equal(e, e0)
    assert g.get_error() == e0[-1]


@pytest.mark.parametrize("alg", ["HIO", "ER", "RAAR"])
def test_split_fields_follow_the_oracle(nd, oracle, synth, alg):
    """Five FP64 iterations at 256^2 through split work fields: estimate and
    every error within 1e-12 of the oracle (x rebuilt from its row transform
    differs from the stored x by rounding only)."""
    d, est, g = job(nd, synth, 256, 44, getattr(nd, alg), True)
    xo = d["x0"].astype(np.complex128).copy()
    o = oracle.OracleCDI(xo, oracle.PLANAR)
    o.set_support(d["support"])
    o.set_intensity(d["intensity"])
    o.set_algorithm(getattr(oracle, alg))
    e = g.iterate(5)
    eo = o.iterate_n(5)
    assert rel(est.get(), xo) <= 1e-12
    for k in range(5):
        assert abs(e[k] - eo[k]) <= 1e-12 * abs(eo[k]), k


def test_split_fields_with_kept_best_estimates(nd, synth):
    """A job keeping its 3 best estimates stores x after every iteration (the
    snapshots read it): the kept estimates and errors are those of the same
    job alone, whether it runs alone in one call or in a batch in pieces."""
    _, est, g = job(nd, synth, 256, 45, nd.HIO, True, n_best=3)
    g.iterate(6)
    ref = [g.get_best_result(k) for k in range(3)]
    _, est2, g2 = job(nd, synth, 256, 45, nd.HIO, True, n_best=3)
    _, _, other = job(nd, synth, 256, 46, nd.HIO, True)
    b = nd.Batch([g2, other])
    b.iterate_each([4, 2])
    b.iterate_each([2, 5])
    for k in range(3):
        xr, er = ref[k]
        x, e = g2.get_best_result(k)
        assert e == er
        assert np.array_equal(x, xr)
