"""nadia_field_magnitude_noalloc through the C ABI: the same magnitudes as
nadia_field_magnitude (bit for bit) on consecutive calls, the second on a
smaller field, staged in the library's fixed device array; fields needing
more staging than it holds are refused with NADIA_ESIZE."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ESIZE = -3


def field(nd, shape, double, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    a[0, 0] = 0  # a zero magnitude
    return nd.Complex2D(*shape, double=double, data=a)


def magnitudes(nd, f, fn):
    out = np.empty((f.nx, f.ny), np.float64 if f.double else np.float32)
    rc = getattr(nd.library(), fn)(f.h, out.ctypes.data_as(ctypes.c_void_p))
    return rc, out


@pytest.mark.parametrize("first,second", [
    (((2048, 2048), True), ((256, 384), True)),     # the largest FP64 field the staging holds, then a small one
    (((4096, 2048), False), ((512, 96), False)),    # the largest FP32 field, then a small one
    (((1024, 1024), False), ((768, 768), True)),    # precision changes between calls
])
def test_noalloc_matches_magnitude(nd, first, second):
    for k, (shape, double) in enumerate((first, second)):
        f = field(nd, shape, double, 90 + k)
        rc, got = magnitudes(nd, f, "nadia_field_magnitude_noalloc")
        assert rc == 0, nd.library().nadia_last_error()
        rc, want = magnitudes(nd, f, "nadia_field_magnitude")
        assert rc == 0
        assert np.array_equal(got, want), (shape, double)
        assert got[0, 0] == 0


def test_noalloc_refuses_larger_fields(nd):
    f = field(nd, (4096, 2048), True, 95)   # 64 MiB of FP64 magnitudes
    rc, _ = magnitudes(nd, f, "nadia_field_magnitude_noalloc")
    assert rc == ESIZE
    assert "nadia_field_magnitude_noalloc" in nd.library().nadia_last_error().decode()
    # a refused call leaves the library usable
    g = field(nd, (128, 128), True, 96)
    rc, got = magnitudes(nd, g, "nadia_field_magnitude_noalloc")
    assert rc == 0
    assert np.array_equal(got, magnitudes(nd, g, "nadia_field_magnitude")[1])
