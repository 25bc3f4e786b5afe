"""Python mirror of NADIA's reconstruction API over the MI355X C ABI.

The product is the C ABI in include/nadia_amd.h (libnadia_hip.so, HIP kernels
for gfx950) and the C++ drop-in classes in include/nadia/.  This module is the
Python view of the same objects, named as in the reference
(NADIA/interfaces/python/pyNADIA/planarcdi.pyx:13-188 forwards the same calls):

    est = Complex2D(nx, ny, double=True)
    planar = PlanarCDI(est)                 # NADIA/include/PlanarCDI.h:24-44
    planar.set_support(support); planar.set_intensity(intensity)
    planar.initialise_estimate(0)
    planar.set_algorithm(HIO)
    planar.iterate(); planar.get_error(); planar.apply_shrinkwrap(1.5, 0.1)

Arrays are numpy, x-major with shape (nx, ny) (element (x, y) at x*ny + y).
There is no CPU fallback: every call runs on the GPU through the C ABI, and
loading fails loudly when the HIP library has not been built.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnadia_hip.so")

FP32, FP64 = 0, 1
PLANAR, FRESNEL, FRESNEL_WF = 0, 1, 2
ER, BIO, BOO, HIO, DM, SF, ASR, HPR, RAAR, CUSTOM = range(10)
PSF, PFS, PS, PF, PI = range(5)
PATH_FUSED, PATH_PLAIN = 0, 1
REAL, IMAG, MAG, PHASE, MAG_SQ = range(5)

_ALG_NAMES = {"ER": ER, "BIO": BIO, "BOO": BOO, "HIO": HIO, "DM": DM, "SF": SF,
              "ASR": ASR, "HPR": HPR, "RAAR": RAAR}

_lib = None


class NadiaError(RuntimeError):
    pass


def _sig(lib):
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    PP = ctypes.POINTER(ctypes.c_void_p)
    IP = ctypes.POINTER(ctypes.c_int)
    DP = ctypes.POINTER(ctypes.c_double)
    LP = ctypes.POINTER(ctypes.c_longlong)
    table = {
        "nadia_last_error": (ctypes.c_char_p, []),
        "nadia_version": (I, []),
        "nadia_device_count": (I, [IP]),
        "nadia_set_device": (I, [I]),
        "nadia_size_supported": (I, [I]),
        "nadia_size_native": (I, [I]),
        "nadia_synchronize": (I, []),
        "nadia_event_create": (I, [PP]),
        "nadia_event_destroy": (I, [P]),
        "nadia_event_record": (I, [P]),
        "nadia_event_query": (I, [P, IP]),
        "nadia_event_synchronize": (I, [P]),
        "nadia_host_alloc": (I, [ctypes.c_size_t, PP]),
        "nadia_host_free": (I, [P]),
        "nadia_recon_get_error_async": (I, [P, P]),
        "nadia_field_create": (I, [I, I, I, PP]),
        "nadia_field_destroy": (I, [P]),
        "nadia_field_dims": (I, [P, IP, IP, IP]),
        "nadia_field_upload": (I, [P, P]),
        "nadia_field_download": (I, [P, P]),
        "nadia_field_copy": (I, [P, P]),
        "nadia_field_fft": (I, [P, I]),
        "nadia_field_invert": (I, [P, I]),
        "nadia_field_scale": (I, [P, D]),
        "nadia_field_add": (I, [P, P, D]),
        "nadia_field_norm": (I, [P, DP]),
        "nadia_field_magnitude": (I, [P, P]),
        "nadia_field_magnitude_noalloc": (I, [P, P]),
        "nadia_recon_create": (I, [P, I, I, PP]),
        "nadia_recon_destroy": (I, [P]),
        "nadia_recon_set_support": (I, [P, P, I]),
        "nadia_recon_set_intensity": (I, [P, P]),
        "nadia_recon_set_beam_stop": (I, [P, P]),
        "nadia_recon_set_algorithm": (I, [P, I]),
        "nadia_recon_set_custom_algorithm": (I, [P, DP]),
        "nadia_recon_set_relaxation_parameter": (I, [P, D]),
        "nadia_recon_get_coefficients": (I, [P, DP, IP]),
        "nadia_recon_iterate": (I, [P, I, DP]),
        "nadia_recon_get_error": (I, [P, DP]),
        "nadia_recon_apply_shrinkwrap": (I, [P, D, D]),
        "nadia_recon_get_support": (I, [P, P]),
        "nadia_recon_initialise_estimate": (I, [P, I]),
        "nadia_recon_apply_support": (I, [P, P]),
        "nadia_recon_project_intensity": (I, [P, P]),
        "nadia_recon_scale_intensity": (I, [P, P]),
        "nadia_recon_propagate_to_detector": (I, [P, P]),
        "nadia_recon_propagate_from_detector": (I, [P, P]),
        "nadia_recon_get_best_result": (I, [P, I, DP, P]),
        "nadia_recon_reset_best": (I, [P]),
        "nadia_recon_get_intensity_autocorrelation": (I, [P, P]),
        "nadia_recon_set_path": (I, [P, I]),
        "nadia_fresnel_setup": (I, [P, P, D, D, D, D, D]),
        "nadia_fresnel_set_experimental_parameters": (I, [P, D, D, D, D]),
        "nadia_fresnel_set_norm": (I, [P, D]),
        "nadia_fresnel_auto_set_norm": (I, [P]),
        "nadia_fresnel_set_normalisation": (I, [P, D]),
        "nadia_fresnel_get_illumination_at_sample": (I, [P, P]),
        "nadia_fresnel_get_transmission_function": (I, [P, P, P]),
        "nadia_fresnel_set_transmission_function": (I, [P, P, P]),
        "nadia_wf_setup": (I, [P, D, D, D, D]),
        "nadia_wf_set_support": (I, [P, D, D]),
        "nadia_batch_create": (I, [PP]),
        "nadia_batch_destroy": (I, [P]),
        "nadia_batch_add": (I, [P, P]),
        "nadia_batch_remove": (I, [P, P]),
        "nadia_batch_size": (I, [P, IP]),
        "nadia_batch_iterate": (I, [P, I, DP]),
        "nadia_batch_iterate_each": (I, [P, IP]),
        "nadia_batch_set_graph": (I, [P, I]),
        "nadia_batch_graph_replays": (I, [P, LP]),
        "nadia_recon_device_bytes": (I, [P, ctypes.POINTER(ctypes.c_size_t)]),
        "nadia_constraint_create": (I, [I, I, PP]),
        "nadia_constraint_destroy": (I, [P]),
        "nadia_constraint_set_charge_flipping": (I, [P, I, I]),
        "nadia_constraint_set_enforce_unity": (I, [P, I]),
        "nadia_constraint_add_region": (I, [P, DP, D, D, IP]),
        "nadia_constraint_set_alphas": (I, [P, I, D, D]),
        "nadia_constraint_set_fixed_c": (I, [P, I, D]),
        "nadia_constraint_get_c_mean": (I, [P, I, DP]),
        "nadia_constraint_clear_regions": (I, [P]),
        "nadia_constraint_set_custom": (I, [P, P, P]),
        "nadia_constraint_apply": (I, [P, P]),
        "nadia_recon_set_complex_constraint": (I, [P, P]),
        "nadia_pd_create": (I, [D, D, I, I, PP]),
        "nadia_pd_destroy": (I, [P]),
        "nadia_pd_add_position": (I, [P, P, D, D, D]),
        "nadia_pd_remove_positions": (I, [P]),
        "nadia_pd_set_parameters": (I, [P, I, D, D]),
        "nadia_pd_set_probe_scaling": (I, [P, I, D]),
        "nadia_pd_initialise_estimate": (I, [P]),
        "nadia_pd_iterate": (I, [P, DP]),
        "nadia_pd_get_size": (I, [P, IP, IP]),
        "nadia_pd_get_transmission": (I, [P, P]),
        "nadia_pd_set_transmission": (I, [P, I, I, P]),
        "nadia_pd_adjust_positions": (I, [P, I, I, I, I, I, I, D, IP]),
        "nadia_pd_get_position": (I, [P, I, DP, DP]),
        "nadia_align": (I, [I, P, I, I, P, I, I, I, I, I, I, P, P, D, IP, IP, IP]),
        "nadia_recon_bytes_estimate": (I, [I, I, I, I, I, ctypes.POINTER(ctypes.c_size_t)]),
        "nadia_recon_set_tile_variants": (I, [P, I, I]),
        "nadia_recon_get_tile_variants": (I, [P, IP, IP]),
        "nadia_tile_variant_count": (I, [IP, IP]),
        "nadia_recon_set_timing": (I, [P, I]),
        "nadia_recon_pass_times": (I, [P, DP, LP]),
    }
    for name, (res, args) in table.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def library():
    """Load libnadia_hip.so (built by __graft_entry__.build()); no fallback."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NadiaError("HIP library %s is not built (run __graft_entry__.build())" % LIB_PATH)
        _lib = _sig(ctypes.CDLL(LIB_PATH))
    return _lib


def _check(rc):
    if rc != 0:
        raise NadiaError(library().nadia_last_error().decode() or ("error %d" % rc))


def _ptr(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(ctypes.c_void_p)


def synchronize():
    _check(library().nadia_synchronize())


class Event:
    """A marker in the library's queue of device work (no counterpart in the
    reference): record() after queuing work, query() / synchronize() later."""

    def __init__(self):
        h = ctypes.c_void_p()
        _check(library().nadia_event_create(ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.nadia_event_destroy(self.h)
            self.h = None

    def record(self):
        _check(library().nadia_event_record(self.h))

    def query(self):
        done = ctypes.c_int(0)
        _check(library().nadia_event_query(self.h, ctypes.byref(done)))
        return bool(done.value)

    def synchronize(self):
        _check(library().nadia_event_synchronize(self.h))


class PinnedDoubles:
    """n doubles of page-locked host memory, as a numpy array (.a): the
    destination of asynchronous reads such as BaseCDI.get_error_async."""

    def __init__(self, n):
        p = ctypes.c_void_p()
        _check(library().nadia_host_alloc(8 * int(n), ctypes.byref(p)))
        self.p = p
        self.a = np.ctypeslib.as_array((ctypes.c_double * int(n)).from_address(p.value))

    def __del__(self):
        if getattr(self, "p", None) and _lib is not None:
            self.a = None
            _lib.nadia_host_free(self.p)
            self.p = None

    def ptr(self, i):
        return ctypes.c_void_p(self.p.value + 8 * int(i))


def get_alg_from_name(name):
    """BaseCDI::getAlgFromName (NADIA/include/BaseCDI.h:382-388): -1 if unknown."""
    return _ALG_NAMES.get(name, -1)


class Complex2D:
    """Device-resident Complex_2D (NADIA/include/Complex_2D.h:35-515)."""

    def __init__(self, nx, ny, double=True, data=None):
        self.nx, self.ny = int(nx), int(ny)
        self.double = bool(double)
        self.cdt = np.complex128 if double else np.complex64
        self.rdt = np.float64 if double else np.float32
        h = ctypes.c_void_p()
        _check(library().nadia_field_create(self.nx, self.ny, FP64 if double else FP32, ctypes.byref(h)))
        self.h = h
        if data is not None:
            self.set(data)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.nadia_field_destroy(self.h)
            self.h = None

    def get_size_x(self):
        return self.nx

    def get_size_y(self):
        return self.ny

    def set(self, a):
        a = np.ascontiguousarray(a, dtype=self.cdt)
        assert a.shape == (self.nx, self.ny)
        _check(library().nadia_field_upload(self.h, _ptr(a)))

    def get(self):
        out = np.empty((self.nx, self.ny), self.cdt)
        _check(library().nadia_field_download(self.h, _ptr(out)))
        return out

    def copy(self, other):
        _check(library().nadia_field_copy(self.h, other.h))

    def perform_forward_fft(self):
        _check(library().nadia_field_fft(self.h, -1))

    def perform_backward_fft(self):
        _check(library().nadia_field_fft(self.h, +1))

    def invert(self, scale=False):
        _check(library().nadia_field_invert(self.h, 1 if scale else 0))

    def scale(self, factor):
        _check(library().nadia_field_scale(self.h, float(factor)))

    def add(self, other, scale=1.0):
        _check(library().nadia_field_add(self.h, other.h, float(scale)))

    def get_norm(self):
        v = ctypes.c_double()
        _check(library().nadia_field_norm(self.h, ctypes.byref(v)))
        return v.value

    def get_2d(self, kind=MAG):
        if kind == MAG:
            out = np.empty((self.nx, self.ny), self.rdt)
            _check(library().nadia_field_magnitude(self.h, _ptr(out)))
            return out
        a = self.get()
        if kind == REAL:
            return a.real.copy()
        if kind == IMAG:
            return a.imag.copy()
        if kind == PHASE:
            return np.angle(a).astype(self.rdt)
        if kind == MAG_SQ:
            return (np.abs(a) ** 2).astype(self.rdt)
        raise ValueError("unknown value type %r" % kind)


class BaseCDI:
    """BaseCDI (NADIA/include/BaseCDI.h:66-522) over nadia_recon."""

    KIND = PLANAR

    def __init__(self, estimate, n_best=0):
        self.complex = estimate
        self.nx, self.ny = estimate.nx, estimate.ny
        self.rdt, self.cdt = estimate.rdt, estimate.cdt
        h = ctypes.c_void_p()
        _check(library().nadia_recon_create(estimate.h, self.KIND, int(n_best), ctypes.byref(h)))
        self.h = h
        self.n_best = n_best

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.nadia_recon_destroy(self.h)
            self.h = None

    def _real(self, a):
        a = np.ascontiguousarray(a, dtype=self.rdt)
        assert a.shape == (self.nx, self.ny), "array dimensions do not match"
        return a

    def get_size_x(self):
        return self.nx

    def get_size_y(self):
        return self.ny

    def set_support(self, support, soften=False):
        _check(library().nadia_recon_set_support(self.h, _ptr(self._real(support)), int(bool(soften))))

    def set_intensity(self, intensity):
        _check(library().nadia_recon_set_intensity(self.h, _ptr(self._real(intensity))))

    def set_beam_stop(self, beam_stop):
        _check(library().nadia_recon_set_beam_stop(self.h, _ptr(self._real(beam_stop))))

    def set_algorithm(self, alg):
        _check(library().nadia_recon_set_algorithm(self.h, int(alg)))

    def set_custom_algorithm(self, *m):
        assert len(m) == 10
        arr = (ctypes.c_double * 10)(*[float(v) for v in m])
        _check(library().nadia_recon_set_custom_algorithm(self.h, arr))

    def set_relaxation_parameter(self, beta):
        _check(library().nadia_recon_set_relaxation_parameter(self.h, float(beta)))

    def get_coefficients(self):
        out = (ctypes.c_double * 5)()
        alg = ctypes.c_int()
        _check(library().nadia_recon_get_coefficients(self.h, out, ctypes.byref(alg)))
        return np.array(list(out)), alg.value

    def iterate(self, n=1):
        """iterate() n times; returns the get_error() value of each iteration."""
        errs = np.zeros(int(n))
        _check(library().nadia_recon_iterate(self.h, int(n), errs.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return errs

    def iterate_async(self, n=1):
        """iterate() n times without copying the errors back (no host sync)."""
        _check(library().nadia_recon_iterate(self.h, int(n), None))

    def get_error(self):
        v = ctypes.c_double()
        _check(library().nadia_recon_get_error(self.h, ctypes.byref(v)))
        return v.value

    def get_error_async(self, pinned, index):
        """Queue a copy of get_error() into pinned.a[index] (PinnedDoubles);
        it is there once an Event recorded after this call has completed."""
        _check(library().nadia_recon_get_error_async(self.h, pinned.ptr(index)))

    def apply_shrinkwrap(self, gauss_width=1.5, threshold=0.1):
        _check(library().nadia_recon_apply_shrinkwrap(self.h, float(gauss_width), float(threshold)))

    def get_support(self):
        out = np.empty((self.nx, self.ny), self.rdt)
        _check(library().nadia_recon_get_support(self.h, _ptr(out)))
        return out

    def initialise_estimate(self, seed=0):
        _check(library().nadia_recon_initialise_estimate(self.h, int(seed)))

    def apply_support(self, c):
        _check(library().nadia_recon_apply_support(self.h, c.h))

    def project_intensity(self, c):
        _check(library().nadia_recon_project_intensity(self.h, c.h))

    def scale_intensity(self, c):
        _check(library().nadia_recon_scale_intensity(self.h, c.h))

    def propagate_to_detector(self, c):
        _check(library().nadia_recon_propagate_to_detector(self.h, c.h))

    def propagate_from_detector(self, c):
        _check(library().nadia_recon_propagate_from_detector(self.h, c.h))

    def set_complex_constraint(self, tc):
        """BaseCDI::set_complex_constraint; None detaches.  The constraint is
        kept alive by this object while attached."""
        self._tc = tc
        _check(library().nadia_recon_set_complex_constraint(self.h, None if tc is None else tc.h))

    def get_best_result(self, index=0):
        out = np.empty((self.nx, self.ny), self.cdt)
        err = ctypes.c_double()
        rc = library().nadia_recon_get_best_result(self.h, int(index), ctypes.byref(err), _ptr(out))
        if rc != 0:
            return None, None
        return out, err.value

    def reset_best(self):
        _check(library().nadia_recon_reset_best(self.h))

    def set_path(self, path):
        _check(library().nadia_recon_set_path(self.h, int(path)))

    def set_tile_variants(self, col_variant, row_variant):
        """Fix the fused passes' tile variants (-1, -1: back to the autotuner)."""
        _check(library().nadia_recon_set_tile_variants(self.h, int(col_variant), int(row_variant)))

    def get_tile_variants(self):
        c, r = ctypes.c_int(), ctypes.c_int()
        _check(library().nadia_recon_get_tile_variants(self.h, ctypes.byref(c), ctypes.byref(r)))
        return c.value, r.value

    def set_timing(self, on=True):
        _check(library().nadia_recon_set_timing(self.h, int(bool(on))))

    def device_bytes(self):
        """Device bytes this reconstruction holds besides its estimate."""
        v = ctypes.c_size_t()
        _check(library().nadia_recon_device_bytes(self.h, ctypes.byref(v)))
        return v.value

    def pass_times(self):
        ms = (ctypes.c_double * 2)()
        cnt = (ctypes.c_longlong * 2)()
        _check(library().nadia_recon_pass_times(self.h, ms, cnt))
        return list(ms), list(cnt)

    @staticmethod
    def getAlgFromName(name):
        return get_alg_from_name(name)


class PlanarCDI(BaseCDI):
    """PlanarCDI (NADIA/include/PlanarCDI.h:24-44)."""

    KIND = PLANAR

    def get_intensity_autocorrelation(self):
        out = np.empty((self.nx, self.ny), self.rdt)
        _check(library().nadia_recon_get_intensity_autocorrelation(self.h, _ptr(out)))
        return out


class FresnelCDI(BaseCDI):
    """FresnelCDI (NADIA/include/FresnelCDI.h:33-232)."""

    KIND = FRESNEL

    def __init__(self, estimate, white_field, beam_wavelength, focal_detector_length,
                 focal_sample_length, pixel_size, normalisation=1.0, n_best=0):
        super().__init__(estimate, n_best)
        wf = np.ascontiguousarray(white_field, dtype=self.cdt)
        assert wf.shape == (self.nx, self.ny)
        _check(library().nadia_fresnel_setup(self.h, _ptr(wf), float(beam_wavelength),
                                             float(focal_detector_length), float(focal_sample_length),
                                             float(pixel_size), float(normalisation)))

    def set_experimental_parameters(self, beam_wavelength, focal_detector_length, focal_sample_length,
                                    pixel_size):
        _check(library().nadia_fresnel_set_experimental_parameters(
            self.h, float(beam_wavelength), float(focal_detector_length), float(focal_sample_length),
            float(pixel_size)))

    def set_norm(self, norm):
        _check(library().nadia_fresnel_set_norm(self.h, float(norm)))

    def auto_set_norm(self):
        _check(library().nadia_fresnel_auto_set_norm(self.h))

    def set_normalisation(self, norm):
        _check(library().nadia_fresnel_set_normalisation(self.h, float(norm)))

    def get_illumination_at_sample(self):
        out = np.empty((self.nx, self.ny), self.cdt)
        _check(library().nadia_fresnel_get_illumination_at_sample(self.h, _ptr(out)))
        return out

    def get_transmission_function(self, esw=None):
        out = np.empty((self.nx, self.ny), self.cdt)
        _check(library().nadia_fresnel_get_transmission_function(self.h, None if esw is None else esw.h,
                                                                 _ptr(out)))
        return out

    def set_transmission_function(self, transmission, esw=None):
        t = np.ascontiguousarray(transmission, dtype=self.cdt)
        _check(library().nadia_fresnel_set_transmission_function(self.h, _ptr(t),
                                                                 None if esw is None else esw.h))


class FresnelCDI_WF(BaseCDI):
    """FresnelCDI_WF (NADIA/include/FresnelCDI_WF.h:31-186)."""

    KIND = FRESNEL_WF

    def __init__(self, estimate, beam_wavelength, zone_focal_length, focal_detector_length, pixel_size,
                 n_best=0):
        super().__init__(estimate, n_best)
        _check(library().nadia_wf_setup(self.h, float(beam_wavelength), float(zone_focal_length),
                                        float(focal_detector_length), float(pixel_size)))

    def set_support(self, z_diameter, size=1.01):  # hides BaseCDI::set_support like the reference
        _check(library().nadia_wf_set_support(self.h, float(z_diameter), float(size)))


class PhaseDiverseCDI:
    """PhaseDiverseCDI (NADIA/include/PhaseDiverseCDI.h) over nadia_pd: frames
    (PlanarCDI / FresnelCDI objects) at positions of one global object."""

    def __init__(self, beta=1.0, gamma=1.0, parallel=False, granularity=1):
        h = ctypes.c_void_p()
        _check(library().nadia_pd_create(float(beta), float(gamma), int(bool(parallel)), int(granularity),
                                         ctypes.byref(h)))
        self.h = h
        self.beta, self.gamma, self.ipc = float(beta), float(gamma), 1
        self.frames = []

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.nadia_pd_destroy(self.h)
            self.h = None

    def add_new_position(self, local, x=0.0, y=0.0, alpha=1.0):
        _check(library().nadia_pd_add_position(self.h, local.h, float(x), float(y), float(alpha)))
        self.frames.append(local)

    def remove_positions(self):
        _check(library().nadia_pd_remove_positions(self.h))
        self.frames = []

    def _params(self):
        _check(library().nadia_pd_set_parameters(self.h, int(self.ipc), self.beta, self.gamma))

    def set_iterations_per_cycle(self, n):
        self.ipc = int(n)
        self._params()

    def set_feedback_parameter(self, beta):
        self.beta = float(beta)
        self._params()

    def set_amplification_factor(self, gamma):
        self.gamma = float(gamma)
        self._params()

    def set_probe_scaling(self, n_probe, alpha):
        _check(library().nadia_pd_set_probe_scaling(self.h, int(n_probe), float(alpha)))

    def initialise_estimate(self):
        _check(library().nadia_pd_initialise_estimate(self.h))

    def iterate(self):
        """One big iteration; returns each frame's get_error()."""
        e = np.zeros(len(self.frames))
        _check(library().nadia_pd_iterate(self.h, e.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return e

    def size(self):
        nx, ny = ctypes.c_int(), ctypes.c_int()
        _check(library().nadia_pd_get_size(self.h, ctypes.byref(nx), ctypes.byref(ny)))
        return nx.value, ny.value

    def get_transmission(self):
        nx, ny = self.size()
        out = np.empty((nx, ny), self.frames[0].cdt)
        _check(library().nadia_pd_get_transmission(self.h, _ptr(out)))
        return out

    def set_transmission(self, t):
        t = np.ascontiguousarray(t, dtype=self.frames[0].cdt)
        _check(library().nadia_pd_set_transmission(self.h, t.shape[0], t.shape[1], _ptr(t)))

    CROSS_CORRELATION, MINIMUM_ERROR = 0, 1

    def adjust_positions(self, type=0, forward=True, x_min=-50, x_max=50, y_min=-50, y_max=50, step_size=4.0):
        """adjust_positions (NADIA/src/PhaseDiverseCDI.c++:337-579); returns the
        per-frame flags of frames whose minimum-error search gave up."""
        gave_up = np.zeros(len(self.frames), np.int32)
        _check(library().nadia_pd_adjust_positions(self.h, int(type), int(bool(forward)), int(x_min), int(x_max),
                                                   int(y_min), int(y_max), float(step_size),
                                                   gave_up.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return gave_up

    def get_position(self, n):
        x, y = ctypes.c_double(), ctypes.c_double()
        _check(library().nadia_pd_get_position(self.h, int(n), ctypes.byref(x), ctypes.byref(y)))
        return x.value, y.value

    def get_final_x_position(self, n):
        return self.get_position(n)[0]

    def get_final_y_position(self, n):
        return self.get_position(n)[1]


def align(img1, img2, min_x=0, max_x=0, min_y=0, max_y=0, w1=None, w2=None, overlap_fraction=0.2, double=True):
    """align (NADIA/src/utils.c:1151-1343) on the device: (found, offset_x,
    offset_y) of img2 against img1."""
    rdt = np.float64 if double else np.float32
    a = np.ascontiguousarray(img1, rdt)
    b = np.ascontiguousarray(img2, rdt)
    wa = None if w1 is None else np.ascontiguousarray(w1, rdt)
    wb = None if w2 is None else np.ascontiguousarray(w2, rdt)
    ox, oy, found = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    _check(library().nadia_align(FP64 if double else FP32, _ptr(a), a.shape[0], a.shape[1], _ptr(b), b.shape[0],
                                 b.shape[1], int(min_x), int(max_x), int(min_y), int(max_y),
                                 None if wa is None else _ptr(wa), None if wb is None else _ptr(wb),
                                 float(overlap_fraction), ctypes.byref(ox), ctypes.byref(oy), ctypes.byref(found)))
    return bool(found.value), ox.value, oy.value


CUSTOM_CONSTRAINT_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_void_p)


class TransmissionConstraint:
    """TransmissionConstraint (NADIA/include/TransmissionConstraint.h) over
    nadia_constraint: unity magnitude and charge flipping on by default;
    add_complex_constraint(region, alpha1, alpha2) adds a ComplexConstraint
    region and returns its index."""

    def __init__(self, nx, ny):
        h = ctypes.c_void_p()
        _check(library().nadia_constraint_create(int(nx), int(ny), ctypes.byref(h)))
        self.h = h
        self.nx, self.ny = nx, ny
        self._custom = None

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.nadia_constraint_destroy(self.h)
            self.h = None

    def set_charge_flipping(self, enable, flip_sign=1):
        _check(library().nadia_constraint_set_charge_flipping(self.h, int(bool(enable)), int(flip_sign)))

    def set_enforce_unity(self, enable):
        _check(library().nadia_constraint_set_enforce_unity(self.h, int(bool(enable))))

    def add_complex_constraint(self, region, alpha1, alpha2, fixed_c=None):
        r = np.ascontiguousarray(region, dtype=np.float64)
        assert r.shape == (self.nx, self.ny)
        k = ctypes.c_int()
        _check(library().nadia_constraint_add_region(self.h, r.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                     float(alpha1), float(alpha2), ctypes.byref(k)))
        if fixed_c is not None:
            _check(library().nadia_constraint_set_fixed_c(self.h, k.value, float(fixed_c)))
        return k.value

    def set_alphas(self, index, alpha1, alpha2):
        _check(library().nadia_constraint_set_alphas(self.h, int(index), float(alpha1), float(alpha2)))

    def c_mean(self, index):
        v = ctypes.c_double()
        _check(library().nadia_constraint_get_c_mean(self.h, int(index), ctypes.byref(v)))
        return v.value

    def delete_complex_constraint_regions(self):
        _check(library().nadia_constraint_clear_regions(self.h))

    def set_custom_constraint(self, fn):
        """fn(array) modifies the transmission function (a numpy view) in place."""
        if fn is None:
            self._custom = None
            _check(library().nadia_constraint_set_custom(self.h, None, None))
            return

        def tramp(ptr, nx, ny, prec, user):
            dt = np.complex128 if prec == FP64 else np.complex64
            buf = (ctypes.c_char * (nx * ny * np.dtype(dt).itemsize)).from_address(ptr)
            fn(np.frombuffer(buf, dtype=dt).reshape(nx, ny))

        self._custom = CUSTOM_CONSTRAINT_FN(tramp)
        _check(library().nadia_constraint_set_custom(self.h, ctypes.cast(self._custom, ctypes.c_void_p), None))

    def apply_constraint(self, c):
        _check(library().nadia_constraint_apply(self.h, c.h))


class Batch:
    """Reconstructions iterated together (nadia_batch_*): jobs of the same
    precision, kind and size share every launch of the fused passes.  Each
    job keeps its own state and gives the same results as alone."""

    def __init__(self, jobs=()):
        h = ctypes.c_void_p()
        _check(library().nadia_batch_create(ctypes.byref(h)))
        self.h = h
        self.jobs = []
        for j in jobs:
            self.add(j)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.nadia_batch_destroy(self.h)
            self.h = None

    def add(self, job):
        _check(library().nadia_batch_add(self.h, job.h))
        self.jobs.append(job)

    def remove(self, job):
        _check(library().nadia_batch_remove(self.h, job.h))
        self.jobs.remove(job)

    def __len__(self):
        return len(self.jobs)

    def iterate(self, n=1):
        """n iterations of every job; returns errors with shape (n, len(self))."""
        errs = np.zeros((int(n), len(self.jobs)))
        if self.jobs:
            _check(library().nadia_batch_iterate(self.h, int(n), errs.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return errs

    def iterate_async(self, n=1):
        _check(library().nadia_batch_iterate(self.h, int(n), None))

    def iterate_each(self, counts):
        """counts[j] iterations of job j (jobs in the order added); queued
        without waiting, like iterate_async."""
        c = np.ascontiguousarray(counts, dtype=np.int32)
        if c.shape != (len(self.jobs),):
            raise ValueError("need one count per job")
        if self.jobs:
            _check(library().nadia_batch_iterate_each(self.h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))

    def set_graph(self, steps):
        """Run uniform iterate() calls as replays of a captured graph of
        `steps` iterations (0: eager launches)."""
        _check(library().nadia_batch_set_graph(self.h, int(steps)))

    def graph_replays(self):
        v = ctypes.c_longlong()
        _check(library().nadia_batch_graph_replays(self.h, ctypes.byref(v)))
        return v.value


def tile_variant_count():
    """(column variants, row variants) of the fused passes."""
    c, r = ctypes.c_int(), ctypes.c_int()
    _check(library().nadia_tile_variant_count(ctypes.byref(c), ctypes.byref(r)))
    return c.value, r.value


def bytes_estimate(nx, ny, double=False, kind=PLANAR, n_best=0):
    """Device bytes a reconstruction of this shape holds, estimate included."""
    v = ctypes.c_size_t()
    _check(library().nadia_recon_bytes_estimate(int(nx), int(ny), FP64 if double else FP32, int(kind), int(n_best),
                                                ctypes.byref(v)))
    return v.value

