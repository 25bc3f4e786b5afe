"""Error behaviour at the boundary: failures return a negative code with a
message in nadia_last_error() (raised as NadiaError by the Python mirror);
the reference's non-failing paths keep their behaviour (an unknown algorithm
number leaves the coefficients alone, getAlgFromName answers -1,
get_best_result answers nothing for a bad index, set_probe_scaling ignores
a bad frame number)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EINVAL, ESIZE, ESTATE = -1, -3, -4


def last_error(nd):
    return nd.library().nadia_last_error().decode()


@pytest.mark.parametrize("shape", [(0, 64), (64, 8192), (48, 5000), (4097, 2), (8, -1)])
def test_unsupported_sizes(nd, shape):
    h = ctypes.c_void_p()
    assert nd.library().nadia_field_create(shape[0], shape[1], nd.FP64, ctypes.byref(h)) == ESIZE
    assert "unsupported size" in last_error(nd)
    with pytest.raises(nd.NadiaError, match="unsupported size"):
        nd.Complex2D(*shape)


def test_bad_arguments(nd):
    lib = nd.library()
    h = ctypes.c_void_p()
    assert lib.nadia_field_create(64, 64, 7, ctypes.byref(h)) == EINVAL
    assert "precision" in last_error(nd)
    f = nd.Complex2D(64, 64)
    r = ctypes.c_void_p()
    assert lib.nadia_recon_create(f.h, 9, 0, ctypes.byref(r)) == EINVAL
    assert lib.nadia_recon_create(f.h, nd.PLANAR, -1, ctypes.byref(r)) == EINVAL
    p = nd.PlanarCDI(f)
    with pytest.raises(nd.NadiaError, match="tile variant out of range"):
        p.set_tile_variants(99, 0)


def test_iterate_needs_intensity(nd):
    p = nd.PlanarCDI(nd.Complex2D(64, 64))
    with pytest.raises(nd.NadiaError, match="intensity not set"):
        p.iterate(1)


def test_constraint_size_mismatch(nd):
    p = nd.PlanarCDI(nd.Complex2D(64, 64))
    with pytest.raises(nd.NadiaError, match="sizes differ"):
        p.set_complex_constraint(nd.TransmissionConstraint(32, 32))


def test_batch_rejects_duplicates(nd):
    p = nd.PlanarCDI(nd.Complex2D(64, 64))
    with pytest.raises(nd.NadiaError, match="already in the batch"):
        nd.Batch([p, p])


def test_reference_non_failing_paths(nd, synth):
    assert nd.get_alg_from_name("no such algorithm") == -1
    d = synth.planar_block(n=64, block=24, support=30)
    est = nd.Complex2D(64, 64, data=d["x0"])
    p = nd.PlanarCDI(est, n_best=2)
    p.set_support(d["support"])
    p.set_intensity(d["intensity"])
    p.set_algorithm(nd.HIO)
    coef = p.get_coefficients()[0]
    p.set_algorithm(99)  # "Algorithm unknown": the coefficients stay
    assert np.array_equal(p.get_coefficients()[0], coef)
    p.iterate(3)
    assert p.get_best_result(5) == (None, None)
    x, e = p.get_best_result(0)
    assert x is not None and e > 0
    pd = nd.PhaseDiverseCDI()
    pd.add_new_position(p)
    pd.set_probe_scaling(3, 0.5)  # out of range: message, no failure
