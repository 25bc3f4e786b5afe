"""Whole-array Complex_2D operations on the device (nadia_field_scale /
add / norm / magnitude; NADIA/src/Complex_2D.c++:169-199, 259-270) against
the oracle's or_scale / or_add / or_get_norm, and the drop-in Complex_2D's
element accessors reading a device-modified array back."""
import numpy as np
import pytest

from test_complex_accessors import PKG, build_probe, expected_elements, run_probe

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("double", [True, False], ids=["fp64", "fp32"])
@pytest.mark.parametrize("shape", [(256, 256), (384, 96)])
def test_field_scale_add_norm_magnitude(nd, oracle, double, shape):
    cdt = np.complex128 if double else np.complex64
    rng = np.random.default_rng(80)
    a = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(cdt)
    b = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(cdt)
    fa = nd.Complex2D(*shape, double=double, data=a)
    fb = nd.Complex2D(*shape, double=double, data=b)
    ref = a.copy()
    fa.scale(-0.375)
    oracle.scale(ref, -0.375, double=double)
    # one rounded product per component on both sides: bit-equal
    assert np.array_equal(fa.get(), ref)
    fa.add(fb, 1.7)
    oracle.add(ref, b, 1.7, double=double)
    tol = 0 if double else 1e-6
    x = fa.get()
    assert np.abs(x - ref).max() <= tol * np.abs(ref).max()
    n_dev, n_ref = fa.get_norm(), oracle.get_norm(ref, double=double)
    # the reference accumulates in the array's real type; the device in double
    assert abs(n_dev - n_ref) <= (1e-13 if double else 1e-4) * n_ref
    assert abs(n_dev - np.linalg.norm(ref.astype(np.complex128))) <= 1e-13 * n_dev
    m = fa.get_2d(nd.MAG)
    assert m.dtype == (np.float64 if double else np.float32)
    assert np.abs(m - np.abs(ref)).max() <= (1e-15 if double else 1e-6) * np.abs(ref).max()


@pytest.mark.parametrize("double", [False, True], ids=["fp32", "fp64"])
def test_cpp_accessors_after_device_ops(double, tmp_path):
    exe = build_probe(tmp_path, double)
    r = run_probe(exe, "device")
    c = np.zeros((8, 24), complex)
    c[:4, :6] = expected_elements()
    c = c * -1.5
    i = np.arange(8)[:, None]
    j = np.arange(24)[None, :]
    d = (0.5 * i - 0.25 * j) + 1j * (1.0 + 0.125 * i * j)
    c = c + 0.75 * d
    tol = 1e-14 if double else 1e-5
    assert len(r["D"]) == 8 * 24
    for ii, jj, re, im, mag in r["D"]:
        want = c[int(ii), int(jj)]
        assert abs(re - want.real) <= tol * 4 and abs(im - want.imag) <= tol * 4
        assert abs(mag - abs(want)) <= tol * 4
    assert abs(r["N"][0][0] - np.linalg.norm(c)) <= tol * np.linalg.norm(c)
