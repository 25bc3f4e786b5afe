"""NADIA's raw binary files (NADIA/src/io_dbin.c:24-100, NADIA/src/io_cplx.c:25-101).

.dbin: nx*ny headerless doubles, y-major (element (x, y) at y*nx + x).
.cplx: the same with interleaved (re, im) pairs.  Arrays here are numpy,
shape (nx, ny) indexed [x, y], the layout of Complex_2D / Double_2D.
"""
import numpy as np


def read_dbin(path, nx, ny):
    a = np.fromfile(path, dtype="<f8")
    if a.size != nx * ny:
        raise ValueError("%s holds %d doubles, not %d x %d" % (path, a.size, nx, ny))
    return np.ascontiguousarray(a.reshape(ny, nx).T)


def write_dbin(path, data):
    np.ascontiguousarray(np.asarray(data, dtype="<f8").T).tofile(path)


def read_cplx(path, nx, ny):
    a = np.fromfile(path, dtype="<f8")
    if a.size != 2 * nx * ny:
        raise ValueError("%s holds %d doubles, not 2 x %d x %d" % (path, a.size, nx, ny))
    a = a.reshape(ny, nx, 2)
    return np.ascontiguousarray((a[..., 0] + 1j * a[..., 1]).T)


def write_cplx(path, data):
    d = np.asarray(data, dtype=np.complex128).T
    np.ascontiguousarray(np.stack([d.real, d.imag], axis=-1)).astype("<f8").tofile(path)
