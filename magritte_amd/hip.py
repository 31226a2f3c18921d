"""ctypes binding of libmagritte_hip.so (include/magritte_hip.h).

``Context`` owns one device-resident model and exposes the reference's
Simulation call sequence: compute_spectral_discretisation,
compute_boundary_intensities, compute_LTE_level_populations,
compute_radiation_field, compute_Jeff, compute_level_populations_from_stateq,
compute_level_populations.  There is no CPU fallback: if the library is not
built or no GPU is present, the calls raise.
"""
import ctypes as C
import os

import numpy as np

from .model import ConvStats

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmagritte_hip.so")

MGR_FP32, MGR_FP64 = 0, 1
MGR_CRIT_FRACTION, MGR_CRIT_MAX = 0, 1

# Entry points declared in include/magritte_hip.h, with (argtypes, restype).
_vp, _i, _i64, _d = C.c_void_p, C.c_int, C.c_int64, C.c_double
SIGNATURES = {
    "mgr_create": ([_vp, _i, _vp], _i),
    "mgr_destroy": ([_vp], None),
    "mgr_last_error": ([_vp], C.c_char_p),
    "mgr_device_count": ([_vp], _i),
    "mgr_compute_spectral_discretisation": ([_vp], _i),
    "mgr_compute_boundary_intensities": ([_vp, _vp], _i),
    "mgr_compute_LTE_level_populations": ([_vp], _i),
    "mgr_trace": ([_vp], _i),
    "mgr_compute_radiation_field": ([_vp, _i], _i),
    "mgr_compute_Jeff": ([_vp], _i),
    "mgr_compute_level_populations_from_stateq": ([_vp, _vp], _i),
    "mgr_iteration_using_Ng_acceleration": ([_vp, _vp], _i),
    "mgr_compute_level_populations": ([_vp, _i, _i64, _i, _i, _vp], _i),
    "mgr_ali_iteration": ([_vp, _i, _i, _vp], _i),
    "mgr_set_populations": ([_vp, _vp], _i),
    "mgr_get_populations": ([_vp, _vp], _i),
    "mgr_get_population_history": ([_vp, _vp, _vp, _vp], _i),
    "mgr_set_population_history": ([_vp, _vp, _vp, _vp], _i),
    "mgr_get_J": ([_vp, _vp], _i),
    "mgr_get_lambda": ([_vp, _vp], _i),
    "mgr_get_Jeff": ([_vp, _vp, _vp, _vp], _i),
    "mgr_get_frequencies": ([_vp, _vp, _vp], _i),
    "mgr_get_boundary_intensities": ([_vp, _vp], _i),
    "mgr_get_emissivity_opacity": ([_vp, _vp, _vp], _i),
    "mgr_get_ray": ([_vp, _i64, _i64, _i64, _vp, _vp, _vp, _vp], _i),
    "mgr_get_npoints_on_rays": ([_vp, _vp], _i),
    "mgr_last_radiation_field_ms": ([_vp, _vp], _i),
    "mgr_trace_stats": ([_vp, _vp, _vp], _i),
}

_LIB = None


class HipError(RuntimeError):
    pass


def load():
    """Load libmagritte_hip.so; raises if it is not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise HipError(f"{LIB_PATH} is not built; run `python -m magritte_amd.build`")
        lib = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = lib
    return _LIB


def device_count():
    n = C.c_int(0)
    load().mgr_device_count(C.byref(n))
    return n.value


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _prec(precision):
    if precision in ("fp32", MGR_FP32, "float32"):
        return MGR_FP32
    if precision in ("fp64", MGR_FP64, "float64"):
        return MGR_FP64
    raise ValueError("precision must be 'fp32' or 'fp64'")


class Context:
    """One device-resident model (mgr_ctx)."""

    def __init__(self, model, device=0):
        self.lib = load()
        self.model = model
        self._view, self._keep = model.view()
        h = C.c_void_p()
        rc = self.lib.mgr_create(C.byref(self._view), device, C.byref(h))
        if rc != 0:
            raise HipError(f"mgr_create failed ({rc}): {self.lib.mgr_last_error(None).decode()}")
        self.h = h
        self.N, self.R, self.F = model.ncells, model.nrays, model.nfreqs
        self.L, self.V, self.Q, self.NB = model.nrad, model.nlev, model.nquads, model.nboundary

    def close(self):
        if getattr(self, "h", None):
            self.lib.mgr_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            raise HipError(f"{what} failed ({rc}): {self.lib.mgr_last_error(self.h).decode()}")

    # ---- set-up -------------------------------------------------------------
    def compute_spectral_discretisation(self):
        self._check(self.lib.mgr_compute_spectral_discretisation(self.h), "compute_spectral_discretisation")

    def compute_boundary_intensities(self, temperatures=None):
        t = None
        if temperatures is not None:
            t = np.ascontiguousarray(temperatures, dtype=np.float64)
        self._check(self.lib.mgr_compute_boundary_intensities(self.h, None if t is None else _p(t)),
                    "compute_boundary_intensities")

    def compute_LTE_level_populations(self):
        self._check(self.lib.mgr_compute_LTE_level_populations(self.h), "compute_LTE_level_populations")

    def setup(self):
        self.compute_spectral_discretisation()
        self.compute_boundary_intensities()
        self.compute_LTE_level_populations()
        return self

    def trace(self):
        self._check(self.lib.mgr_trace(self.h), "trace")

    # ---- hot path -----------------------------------------------------------
    def compute_radiation_field(self, precision="fp32"):
        self._check(self.lib.mgr_compute_radiation_field(self.h, _prec(precision)), "compute_radiation_field")

    def compute_Jeff(self):
        self._check(self.lib.mgr_compute_Jeff(self.h), "compute_Jeff")

    def compute_level_populations_from_stateq(self):
        c = ConvStats()
        self._check(self.lib.mgr_compute_level_populations_from_stateq(self.h, C.byref(c)),
                    "compute_level_populations_from_stateq")
        return c.fraction_not_converged, c.relative_change_mean, c.relative_change_max

    def iteration_using_Ng_acceleration(self):
        c = ConvStats()
        self._check(self.lib.mgr_iteration_using_Ng_acceleration(self.h, C.byref(c)),
                    "iteration_using_Ng_acceleration")
        return c.fraction_not_converged, c.relative_change_mean, c.relative_change_max

    def compute_level_populations(self, use_Ng=True, max_niterations=1000, criterion="fraction",
                                  precision="fp32"):
        crit = {"fraction": MGR_CRIT_FRACTION, "max": MGR_CRIT_MAX}.get(criterion, criterion)
        it = C.c_int64(-1)
        self._check(self.lib.mgr_compute_level_populations(self.h, 1 if use_Ng else 0, max_niterations,
                                                           crit, _prec(precision), C.byref(it)),
                    "compute_level_populations")
        return it.value

    def ali_iteration(self, ng=False, precision="fp32"):
        c = ConvStats()
        self._check(self.lib.mgr_ali_iteration(self.h, 1 if ng else 0, _prec(precision), C.byref(c)),
                    "ali_iteration")
        return c.fraction_not_converged, c.relative_change_mean, c.relative_change_max

    # ---- data ---------------------------------------------------------------
    def J(self):
        a = np.zeros(self.N * self.F)
        self._check(self.lib.mgr_get_J(self.h, _p(a)), "get_J")
        return a.reshape(self.N, self.F)

    def lambda_diag(self):
        a = np.zeros(self.N * self.L)
        self._check(self.lib.mgr_get_lambda(self.h, _p(a)), "get_lambda")
        return a.reshape(self.N, self.L)

    def Jeff(self):
        a, b, c = (np.zeros(self.N * self.L) for _ in range(3))
        self._check(self.lib.mgr_get_Jeff(self.h, _p(a), _p(b), _p(c)), "get_Jeff")
        return a.reshape(self.N, self.L), b.reshape(self.N, self.L), c.reshape(self.N, self.L)

    def populations(self):
        a = np.zeros(self.N * self.V)
        self._check(self.lib.mgr_get_populations(self.h, _p(a)), "get_populations")
        return a.reshape(self.N, self.V)

    def set_populations(self, pops):
        a = np.ascontiguousarray(pops, dtype=np.float64).ravel()
        self._check(self.lib.mgr_set_populations(self.h, _p(a)), "set_populations")

    def population_history(self):
        a, b, c = (np.zeros(self.N * self.V) for _ in range(3))
        self._check(self.lib.mgr_get_population_history(self.h, _p(a), _p(b), _p(c)), "get_population_history")
        return tuple(x.reshape(self.N, self.V) for x in (a, b, c))

    def set_population_history(self, p1, p2, p3):
        arrs = [np.ascontiguousarray(x, dtype=np.float64).ravel() for x in (p1, p2, p3)]
        self._check(self.lib.mgr_set_population_history(self.h, *[_p(x) for x in arrs]),
                    "set_population_history")

    def frequencies(self):
        nu = np.zeros(self.N * self.F)
        nr = np.zeros(self.N * self.L * self.Q, np.int64)
        self._check(self.lib.mgr_get_frequencies(self.h, _p(nu), _p(nr)), "get_frequencies")
        return nu.reshape(self.N, self.F), nr.reshape(self.N, self.L, self.Q)

    def boundary_intensities(self):
        a = np.zeros(self.NB * self.F)
        self._check(self.lib.mgr_get_boundary_intensities(self.h, _p(a)), "get_boundary_intensities")
        return a.reshape(self.NB, self.F)

    def emissivity_opacity(self):
        a, b = np.zeros(self.N * self.L), np.zeros(self.N * self.L)
        self._check(self.lib.mgr_get_emissivity_opacity(self.h, _p(a), _p(b)), "get_emissivity_opacity")
        return a.reshape(self.N, self.L), b.reshape(self.N, self.L)

    def ray(self, origin, ray):
        n = C.c_int64(0)
        self._check(self.lib.mgr_get_ray(self.h, origin, ray, 0, None, None, None, C.byref(n)), "get_ray")
        cap = max(n.value, 1)
        cell = np.zeros(cap, np.int64); sh = np.zeros(cap); dz = np.zeros(cap)
        self._check(self.lib.mgr_get_ray(self.h, origin, ray, cap, _p(cell), _p(sh), _p(dz), C.byref(n)),
                    "get_ray")
        k = n.value
        return cell[:k], sh[:k], dz[:k]

    def npoints_on_rays(self):
        a = np.zeros((self.R // 2) * self.N, np.int64)
        self._check(self.lib.mgr_get_npoints_on_rays(self.h, _p(a)), "get_npoints_on_rays")
        return a.reshape(self.R // 2, self.N)

    def last_radiation_field_ms(self):
        v = C.c_double(0.0)
        self._check(self.lib.mgr_last_radiation_field_ms(self.h, C.byref(v)), "last_radiation_field_ms")
        return v.value

    def trace_stats(self):
        a, b = C.c_int64(0), C.c_int64(0)
        self._check(self.lib.mgr_trace_stats(self.h, C.byref(a), C.byref(b)), "trace_stats")
        return a.value, b.value
