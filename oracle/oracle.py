"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  Arrays are numpy, x-major (shape
(nx, ny), element (x, y) at x*ny + y), the layout of NADIA's Complex_2D
(NADIA/include/Complex_2D.h:179-194) and Double_2D (NADIA/include/Double_2D.h:121-143).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

ER, BIO, BOO, HIO, DM, SF, ASR, HPR, RAAR, CUSTOM = range(10)
PSF, PFS, PS, PF, PI = range(5)
PLANAR, FRESNEL, FRESNEL_WF = 0, 1, 2

_libs = {}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _lib(double=True, timing=False):
    """The oracle library; timing=True selects the CPU-baseline builds
    (-O3 -ffast-math; FP32 with threaded 2-D transforms), never the checker."""
    key = ("d" if double else "f") + ("t" if timing else "")
    if key in _libs:
        return _libs[key]
    path = os.path.join(HERE, "liboracle_%s.so" % key)
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    sig = {
        "or_fft2d": (None, [P, I, I, I]),
        "or_invert": (None, [P, I, I, I]),
        "ocdi_create": (P, [P, I, I, I, I]),
        "ocdi_destroy": (None, [P]),
        "ocdi_convolve": (None, [P, P, D, I]),
        "ocdi_set_support": (None, [P, P, I]),
        "ocdi_set_beam_stop": (None, [P, P]),
        "ocdi_set_intensity": (None, [P, P]),
        "ocdi_get_error": (D, [P]),
        "ocdi_set_transmission_function": (None, [P, P, P]),
        "ocdi_get_transmission_function": (None, [P, P, P]),
        "ocdi_propagate_to_detector": (None, [P, P]),
        "ocdi_propagate_from_detector": (None, [P, P]),
        "ocdi_scale_intensity": (None, [P, P]),
        "ocdi_apply_support": (None, [P, P]),
        "ocdi_project_intensity": (None, [P, P]),
        "ocdi_set_custom_algorithm": (None, [P] + [D] * 10),
        "ocdi_set_algorithm": (None, [P, I]),
        "ocdi_set_relaxation_parameter": (None, [P, D]),
        "ocdi_get_coefficients": (None, [P, P, P]),
        "ocdi_get_best_result": (I, [P, I, P, P]),
        "ocdi_reset_best": (None, [P]),
        "ocdi_iterate": (I, [P]),
        "ocdi_iterate_n": (None, [P, I, P]),
        "ocdi_apply_shrinkwrap": (None, [P, D, D]),
        "ocdi_get_support": (None, [P, P]),
        "ocdi_get_intensity_sqrt": (None, [P, P]),
        "ocdi_initialise_estimate": (None, [P, I]),
        "ocdi_get_intensity_autocorrelation": (None, [P, P]),
        "ocdi_fresnel_init": (None, [P, P, D, D, D, D, D]),
        "ocdi_fresnel_set_experimental_parameters": (None, [P, D, D, D, D]),
        "ocdi_fresnel_set_norm": (None, [P, D]),
        "ocdi_fresnel_auto_set_norm": (None, [P]),
        "ocdi_fresnel_get_illumination_at_sample": (None, [P, P]),
        "ocdi_fresnel_get_illumination": (None, [P, P]),
        "ocdi_wf_init": (None, [P, D, D, D, D]),
        "ocdi_wf_set_support": (None, [P, D, D]),
        "oracle_real_size": (I, []),
        "otc_create": (P, [I, I]),
        "otc_destroy": (None, [P]),
        "otc_set_charge_flipping": (None, [P, I, I]),
        "otc_set_enforce_unity": (None, [P, I]),
        "otc_add_region": (I, [P, P, D, D]),
        "otc_set_fixed_c": (None, [P, I, D]),
        "otc_get_c_mean": (D, [P, I]),
        "otc_apply": (None, [P, P]),
        "ocdi_set_complex_constraint": (None, [P, P]),
        "opd_create": (P, [D, D, I, I]),
        "opd_destroy": (None, [P]),
        "opd_set_iterations_per_cycle": (None, [P, I]),
        "opd_set_feedback_parameter": (None, [P, D]),
        "opd_set_amplification_factor": (None, [P, D]),
        "opd_set_probe_scaling": (None, [P, I, D]),
        "opd_add_position": (I, [P, P, D, D, D]),
        "opd_initialise_estimate": (None, [P]),
        "opd_iterate": (None, [P]),
        "opd_get_size": (None, [P, P, P]),
        "opd_get_transmission": (None, [P, P]),
        "opd_set_transmission": (I, [P, P, I, I]),
        "opd_get_weight": (None, [P, I, P]),
        "opd_get_origin": (None, [P, P, P]),
        "opd_adjust_positions": (I, [P, I, I, I, I, I, I, D, P]),
        "opd_get_position": (None, [P, I, P, P]),
        "oalign": (I, [P, I, I, P, I, I, P, P, I, I, I, I, P, P, D]),
        "ointerpolate": (None, [P, I, I, P, I, I]),
    }
    R = D if double else ctypes.c_float  # the oracle's `real`
    S = ctypes.c_size_t
    sig.update({"or_add": (None, [P, P, S, R]), "or_scale": (None, [P, S, R]), "or_get_norm": (R, [P, S])})
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _libs[key] = lib
    return lib


def _ptr(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def fft2d(a, sign, double=True):
    """Unnormalised in-place 2-D DFT (FFTW convention: sign -1 forward)."""
    lib = _lib(double)
    lib.or_fft2d(_ptr(a), a.shape[0], a.shape[1], sign)
    return a


def invert(a, scale=True, double=True):
    """Complex_2D::invert (fftshift with optional 1/sqrt(N) scale), in place."""
    lib = _lib(double)
    lib.or_invert(_ptr(a), a.shape[0], a.shape[1], 1 if scale else 0)
    return a


def add(a, b, scale=1.0, double=True):
    """Complex_2D::add (a += scale * b), in place."""
    _lib(double).or_add(_ptr(a), _ptr(b), a.size, scale)
    return a


def scale(a, factor, double=True):
    """Complex_2D::scale, in place."""
    _lib(double).or_scale(_ptr(a), a.size, factor)
    return a


def get_norm(a, double=True):
    """Complex_2D::get_norm: sqrt of the sum of |c|^2, accumulated in the
    array's real type."""
    return _lib(double).or_get_norm(_ptr(a), a.size)


class OracleConstraint:
    """TransmissionConstraint (NADIA/src/TransmissionConstraint.c++:26-172)."""

    def __init__(self, nx, ny, double=True):
        self.double = double
        self.lib = _lib(double)
        self.rdt = np.float64 if double else np.float32
        self.cdt = np.complex128 if double else np.complex64
        self.h = self.lib.otc_create(nx, ny)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.otc_destroy(self.h)
            self.h = None

    def set_charge_flipping(self, enable, flip_sign=1):
        self.lib.otc_set_charge_flipping(self.h, int(enable), int(flip_sign))

    def set_enforce_unity(self, enable):
        self.lib.otc_set_enforce_unity(self.h, int(enable))

    def add_complex_constraint(self, region, alpha1, alpha2, fixed_c=None):
        r = np.ascontiguousarray(region, dtype=self.rdt)
        k = self.lib.otc_add_region(self.h, _ptr(r), alpha1, alpha2)
        if fixed_c is not None:
            self.lib.otc_set_fixed_c(self.h, k, fixed_c)
        return k

    def c_mean(self, k):
        return self.lib.otc_get_c_mean(self.h, k)

    def apply(self, t):
        """apply_constraint on a complex array, in place."""
        assert t.dtype == self.cdt
        self.lib.otc_apply(self.h, _ptr(t))
        return t


class OracleCDI:
    """One BaseCDI-derived object (PlanarCDI, FresnelCDI or FresnelCDI_WF)."""

    def __init__(self, estimate, kind=PLANAR, n_best=0, double=True, timing=False):
        self.double = double
        self.cdt = np.complex128 if double else np.complex64
        self.rdt = np.float64 if double else np.float32
        assert estimate.dtype == self.cdt and estimate.ndim == 2
        self.lib = _lib(double, timing)
        self.nx, self.ny = estimate.shape
        self.complex = estimate  # mutated in place, like BaseCDI::complex
        self.h = self.lib.ocdi_create(_ptr(estimate), self.nx, self.ny, kind, n_best)
        self.kind = kind
        self.n_best = n_best

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ocdi_destroy(self.h)
            self.h = None

    def _r(self, a):
        return np.ascontiguousarray(a, dtype=self.rdt)

    def _c(self, a):
        return np.ascontiguousarray(a, dtype=self.cdt)

    # -- BaseCDI ----------------------------------------------------------
    def set_support(self, s, soften=False):
        s = self._r(s)
        self.lib.ocdi_set_support(self.h, _ptr(s), int(soften))

    def set_intensity(self, i):
        i = self._r(i)
        self.lib.ocdi_set_intensity(self.h, _ptr(i))

    def set_beam_stop(self, b):
        b = self._r(b)
        self.lib.ocdi_set_beam_stop(self.h, _ptr(b))

    def set_algorithm(self, alg):
        self.lib.ocdi_set_algorithm(self.h, alg)

    def set_custom_algorithm(self, *m):
        assert len(m) == 10
        self.lib.ocdi_set_custom_algorithm(self.h, *[float(v) for v in m])

    def set_relaxation_parameter(self, beta):
        self.lib.ocdi_set_relaxation_parameter(self.h, beta)

    def coefficients(self):
        out = np.zeros(5)
        alg = ctypes.c_int(0)
        self.lib.ocdi_get_coefficients(self.h, _ptr(out), ctypes.byref(alg))
        return out, alg.value

    def iterate(self):
        return self.lib.ocdi_iterate(self.h)

    def iterate_n(self, n):
        errs = np.zeros(n)
        self.lib.ocdi_iterate_n(self.h, n, _ptr(errs))
        return errs

    def get_error(self):
        return self.lib.ocdi_get_error(self.h)

    def apply_shrinkwrap(self, gauss_width=1.5, threshold=0.1):
        self.lib.ocdi_apply_shrinkwrap(self.h, gauss_width, threshold)

    def get_support(self):
        out = np.zeros((self.nx, self.ny), self.rdt)
        self.lib.ocdi_get_support(self.h, _ptr(out))
        return out

    def get_intensity_sqrt(self):
        out = np.zeros((self.nx, self.ny), self.rdt)
        self.lib.ocdi_get_intensity_sqrt(self.h, _ptr(out))
        return out

    def convolve(self, arr, gauss_width, cutoff=4):
        a = self._r(arr).copy()
        self.lib.ocdi_convolve(self.h, _ptr(a), gauss_width, cutoff)
        return a

    def initialise_estimate(self, seed=0):
        self.lib.ocdi_initialise_estimate(self.h, seed)

    def apply_support(self, c):
        self.lib.ocdi_apply_support(self.h, _ptr(c))

    def scale_intensity(self, c):
        self.lib.ocdi_scale_intensity(self.h, _ptr(c))

    def project_intensity(self, c):
        self.lib.ocdi_project_intensity(self.h, _ptr(c))

    def propagate_to_detector(self, c):
        self.lib.ocdi_propagate_to_detector(self.h, _ptr(c))

    def propagate_from_detector(self, c):
        self.lib.ocdi_propagate_from_detector(self.h, _ptr(c))

    def set_complex_constraint(self, tc):
        self.tc = tc  # keep it alive: the object holds a non-owning pointer
        self.lib.ocdi_set_complex_constraint(self.h, None if tc is None else tc.h)

    def get_best_result(self, index=0):
        out = np.zeros((self.nx, self.ny), self.cdt)
        err = ctypes.c_double(0)
        ok = self.lib.ocdi_get_best_result(self.h, index, ctypes.byref(err), _ptr(out))
        return (out, err.value) if ok else (None, None)

    def reset_best(self):
        self.lib.ocdi_reset_best(self.h)

    def get_intensity_autocorrelation(self):
        out = np.zeros((self.nx, self.ny), self.rdt)
        self.lib.ocdi_get_intensity_autocorrelation(self.h, _ptr(out))
        return out

    # -- FresnelCDI -------------------------------------------------------
    def fresnel_init(self, white_field, wavelength, zfd, zfs, pixel, norm=1.0):
        wf = self._c(white_field)
        self.lib.ocdi_fresnel_init(self.h, _ptr(wf), wavelength, zfd, zfs, pixel, norm)

    def set_norm(self, n):
        self.lib.ocdi_fresnel_set_norm(self.h, n)

    def auto_set_norm(self):
        self.lib.ocdi_fresnel_auto_set_norm(self.h)

    def get_illumination_at_sample(self):
        out = np.zeros((self.nx, self.ny), self.cdt)
        self.lib.ocdi_fresnel_get_illumination_at_sample(self.h, _ptr(out))
        return out

    def get_illumination(self):
        out = np.zeros((self.nx, self.ny), self.cdt)
        self.lib.ocdi_fresnel_get_illumination(self.h, _ptr(out))
        return out

    def get_transmission_function(self, esw=None):
        out = np.zeros((self.nx, self.ny), self.cdt)
        e = None if esw is None else self._c(esw)
        self.lib.ocdi_get_transmission_function(self.h, _ptr(out), None if e is None else _ptr(e))
        return out

    def set_transmission_function(self, T, esw=None):
        T = self._c(T)
        self.lib.ocdi_set_transmission_function(self.h, _ptr(T), None if esw is None else _ptr(esw))

    # -- FresnelCDI_WF ----------------------------------------------------
    def wf_init(self, wavelength, z12, z23, pixel):
        self.lib.ocdi_wf_init(self.h, wavelength, z12, z23, pixel)

    def wf_set_support(self, z_diameter, size=1.01):
        self.lib.ocdi_wf_set_support(self.h, z_diameter, size)


class OraclePhaseDiverse:
    """PhaseDiverseCDI (NADIA/src/PhaseDiverseCDI.c++) over OracleCDI frames."""

    def __init__(self, beta=1.0, gamma=1.0, parallel=False, granularity=1, double=True):
        self.double = double
        self.lib = _lib(double)
        self.cdt = np.complex128 if double else np.complex64
        self.rdt = np.float64 if double else np.float32
        self.h = self.lib.opd_create(beta, gamma, int(parallel), int(granularity))
        self.frames = []

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.opd_destroy(self.h)
            self.h = None

    def add_new_position(self, frame, x=0.0, y=0.0, alpha=1.0):
        self.frames.append(frame)  # the frames must outlive this object
        return self.lib.opd_add_position(self.h, frame.h, float(x), float(y), float(alpha))

    def set_iterations_per_cycle(self, n):
        self.lib.opd_set_iterations_per_cycle(self.h, int(n))

    def set_feedback_parameter(self, beta):
        self.lib.opd_set_feedback_parameter(self.h, float(beta))

    def set_amplification_factor(self, gamma):
        self.lib.opd_set_amplification_factor(self.h, float(gamma))

    def set_probe_scaling(self, n, alpha):
        self.lib.opd_set_probe_scaling(self.h, int(n), float(alpha))

    def initialise_estimate(self):
        self.lib.opd_initialise_estimate(self.h)

    def iterate(self):
        self.lib.opd_iterate(self.h)

    def size(self):
        nx, ny = ctypes.c_int(), ctypes.c_int()
        self.lib.opd_get_size(self.h, ctypes.byref(nx), ctypes.byref(ny))
        return nx.value, ny.value

    def origin(self):
        x, y = ctypes.c_int(), ctypes.c_int()
        self.lib.opd_get_origin(self.h, ctypes.byref(x), ctypes.byref(y))
        return x.value, y.value

    def get_transmission(self):
        nx, ny = self.size()
        out = np.zeros((nx, ny), self.cdt)
        self.lib.opd_get_transmission(self.h, _ptr(out))
        return out

    def set_transmission(self, t):
        t = np.ascontiguousarray(t, dtype=self.cdt)
        return self.lib.opd_set_transmission(self.h, _ptr(t), t.shape[0], t.shape[1])

    def get_weight(self, n):
        f = self.frames[n]
        out = np.zeros((f.nx, f.ny), self.rdt)
        self.lib.opd_get_weight(self.h, int(n), _ptr(out))
        return out

    CROSS_CORRELATION, MINIMUM_ERROR = 0, 1

    def adjust_positions(self, type=0, forward=True, x_min=-50, x_max=50, y_min=-50, y_max=50, step_size=4.0):
        """Returns the per-frame 'gave up' flags of the minimum-error search."""
        gave_up = np.zeros(len(self.frames), np.int32)
        rc = self.lib.opd_adjust_positions(self.h, int(type), int(bool(forward)), int(x_min), int(x_max), int(y_min),
                                           int(y_max), float(step_size), _ptr(gave_up))
        if rc != 0:
            raise ValueError("adjust_positions: unknown type or frames of different sizes")
        return gave_up

    def get_position(self, n):
        x, y = ctypes.c_double(), ctypes.c_double()
        self.lib.opd_get_position(self.h, int(n), ctypes.byref(x), ctypes.byref(y))
        return x.value, y.value


def oracle_align(a, b, min_x=0, max_x=0, min_y=0, max_y=0, wa=None, wb=None, overlap_fraction=0.2, double=True):
    """align (NADIA/src/utils.c:1151-1343): returns (found, offset_x, offset_y)."""
    lib = _lib(double)
    rdt = np.float64 if double else np.float32
    a = np.ascontiguousarray(a, rdt)
    b = np.ascontiguousarray(b, rdt)
    wa = None if wa is None else np.ascontiguousarray(wa, rdt)
    wb = None if wb is None else np.ascontiguousarray(wb, rdt)
    ox, oy = ctypes.c_int(0), ctypes.c_int(0)
    found = lib.oalign(_ptr(a), a.shape[0], a.shape[1], _ptr(b), b.shape[0], b.shape[1], ctypes.byref(ox),
                       ctypes.byref(oy), int(min_x), int(max_x), int(min_y), int(max_y),
                       None if wa is None else _ptr(wa), None if wb is None else _ptr(wb), float(overlap_fraction))
    return bool(found), ox.value, oy.value


def oracle_interpolate(a, scale, double=True):
    """interpolate(Double_2D) (NADIA/src/utils.c:1346-1435) by an integer factor."""
    lib = _lib(double)
    rdt = np.float64 if double else np.float32
    a = np.ascontiguousarray(a, rdt)
    out = np.zeros((a.shape[0] * scale, a.shape[1] * scale), rdt)
    lib.ointerpolate(_ptr(a), a.shape[0], a.shape[1], _ptr(out), out.shape[0], out.shape[1])
    return out
