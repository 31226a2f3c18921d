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
    "mgr_create_batch": ([_vp, _i64, _i, _vp], _i),
    "mgr_batch_info": ([_vp, _vp, _vp], _i),
    "mgr_batch_convergence": ([_vp, _vp, _vp], _i),
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
    "mgr_lambda_band_slots": ([_vp, _vp], _i),
    "mgr_get_lambda_band": ([_vp, _vp, _vp], _i),
    "mgr_band_sweeps": ([_vp, _vp, _vp], _i),
    "mgr_species_info": ([_vp, _vp, _vp, _vp], _i),
    "mgr_species_convergence": ([_vp, _i64, _vp], _i),
    "mgr_get_Jeff": ([_vp, _vp, _vp, _vp], _i),
    "mgr_get_frequencies": ([_vp, _vp, _vp], _i),
    "mgr_get_boundary_intensities": ([_vp, _vp], _i),
    "mgr_get_emissivity_opacity": ([_vp, _vp, _vp], _i),
    "mgr_get_ray": ([_vp, _i64, _i64, _i64, _vp, _vp, _vp, _vp], _i),
    "mgr_get_npoints_on_rays": ([_vp, _vp], _i),
    "mgr_get_npoints_on_rays_rest": ([_vp, _vp], _i),
    "mgr_set_scattering": ([_vp, _i], _i),
    "mgr_set_window_skip": ([_vp, _i], _i),
    "mgr_set_far_cache": ([_vp, _i], _i),
    "mgr_set_trace_options": ([_vp, _i64, _i], _i),
    "mgr_window_skip_stats": ([_vp, _vp, _vp], _i),
    "mgr_get_u": ([_vp, _vp], _i),
    "mgr_compute_spectral_discretisation_image": ([_vp, _d], _i),
    "mgr_compute_image": ([_vp, _i64, _i, _vp, _vp], _i),
    "mgr_last_radiation_field_ms": ([_vp, _vp], _i),
    "mgr_last_radiation_field_kernel": ([_vp, C.c_char_p, _i], _i),
    "mgr_first_pass": ([_vp, _i, _i], _i),
    "mgr_prepare": ([_vp, _i], _i),
    "mgr_ng_acceleration_noalloc": ([_i64, _i64, C.c_double, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "mgr_trace_stats": ([_vp, _vp, _vp], _i),
    "mgr_radiation_field_shape": ([_vp, _i, _vp, _vp, _i, _vp, _vp, _vp, _vp], _i),
    "mgr_set_radiation_field_shape": ([_vp, _i, _i, _i], _i),
    "mgr_mlp_num_params": ([], _i64),
    "mgr_mlp_create": ([_i, C.c_uint64, _vp], _i),
    "mgr_mlp_destroy": ([_vp], None),
    "mgr_mlp_last_error": ([_vp], C.c_char_p),
    "mgr_mlp_set_params": ([_vp, _vp], _i),
    "mgr_mlp_get_params": ([_vp, _vp], _i),
    "mgr_mlp_set_normalisation": ([_vp, _vp, _vp, _vp, _vp], _i),
    "mgr_mlp_forward": ([_vp, _vp, _i64, _vp], _i),
    "mgr_mlp_loss_grad": ([_vp, _vp, _vp, _i64, _vp, _vp], _i),
    "mgr_mlp_train": ([_vp, _vp, _vp, _i64, _i64, _i64, C.c_float, C.c_uint64, _vp], _i),
    "mgr_features": ([_vp, _vp], _i),
    "mgr_warm_start": ([_vp, _vp], _i),
    "mgr_warm_start_fallbacks": ([_vp, _vp], _i),
    "mgr_mlp_set_jlin_inputs": ([_vp, _i], _i),
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
        self.L, self.V, self.Q, self.NB = model.nlines, model.nlev, model.nquads, model.nboundary
        # levels of each line-producing species; populations hold them species after species
        self.Vs = [ld.nlev for ld in model.species_linedata]

    def _split(self, a):
        """A population array as (N, V) for one species, else a list per species."""
        if len(self.Vs) == 1:
            return a.reshape(self.N, self.V)
        out, o = [], 0
        for V in self.Vs:
            out.append(a[o:o + self.N * V].reshape(self.N, V))
            o += self.N * V
        return out

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

    def compute_spectral_discretisation_image(self, width):
        self._check(self.lib.mgr_compute_spectral_discretisation_image(self.h, float(width)),
                    "compute_spectral_discretisation_image")

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

    def lambda_band_raw(self):
        """Band slots of Lambda* (n_off_diag >= 2): cells [N, S] (-1: no
        element) and Ls [N, S, L]."""
        S = C.c_int64()
        self._check(self.lib.mgr_lambda_band_slots(self.h, C.byref(S)), "lambda_band_slots")
        S = S.value
        cells = np.zeros(self.N * S, dtype=np.int32)
        Ls = np.zeros(self.N * S * self.L)
        self._check(self.lib.mgr_get_lambda_band(self.h, _p(cells), _p(Ls)), "get_lambda_band")
        return cells.reshape(self.N, S), Ls.reshape(self.N, S, self.L)

    def species_convergence(self):
        """(fraction not converged, mean change, max change) of every species
        after the last level-population update."""
        out = []
        for l in range(len(self.Vs)):
            c = ConvStats()
            self._check(self.lib.mgr_species_convergence(self.h, l, C.byref(c)), "species_convergence")
            out.append((c.fraction_not_converged, c.relative_change_mean, c.relative_change_max))
        return out

    def band_sweeps(self):
        """Block-Jacobi sweeps of the last rate-equation solve with bands and
        the largest population move of each (over the cell's total)."""
        n = C.c_int()
        ch = np.zeros(16)
        self._check(self.lib.mgr_band_sweeps(self.h, C.byref(n), _p(ch)), "band_sweeps")
        return n.value, ch[:n.value]

    def lambda_band(self):
        """Band elements merged per (p, k) and emitting cell, as
        Lambda::add_element keeps them: {(p, k): {nr: Ls}} for nr != p; an
        element naming p itself is added to lambda_diag()'s (returned
        separately as the second value)."""
        cells, Ls = self.lambda_band_raw()
        out, diag_add = {}, np.zeros((self.N, self.L))
        for p, s in zip(*np.nonzero(cells >= 0)):
            c = int(cells[p, s])
            for k in range(self.L):
                if c == p:
                    diag_add[p, k] += Ls[p, s, k]
                else:
                    d = out.setdefault((int(p), k), {})
                    d[c] = d.get(c, 0.0) + Ls[p, s, k]
        return out, diag_add

    def Jeff(self):
        a, b, c = (np.zeros(self.N * self.L) for _ in range(3))
        self._check(self.lib.mgr_get_Jeff(self.h, _p(a), _p(b), _p(c)), "get_Jeff")
        return a.reshape(self.N, self.L), b.reshape(self.N, self.L), c.reshape(self.N, self.L)

    def populations(self):
        a = np.zeros(self.N * sum(self.Vs))
        self._check(self.lib.mgr_get_populations(self.h, _p(a)), "get_populations")
        return self._split(a)

    def set_populations(self, pops):
        if isinstance(pops, (list, tuple)):
            pops = np.concatenate([np.ravel(x) for x in pops])
        a = np.ascontiguousarray(pops, dtype=np.float64).ravel()
        self._check(self.lib.mgr_set_populations(self.h, _p(a)), "set_populations")

    def population_history(self):
        a, b, c = (np.zeros(self.N * sum(self.Vs)) for _ in range(3))
        self._check(self.lib.mgr_get_population_history(self.h, _p(a), _p(b), _p(c)), "get_population_history")
        return tuple(self._split(x) for x in (a, b, c))

    def set_population_history(self, p1, p2, p3):
        p1, p2, p3 = (np.concatenate([np.ravel(y) for y in x]) if isinstance(x, (list, tuple)) else x
                      for x in (p1, p2, p3))
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

    def set_scattering(self, use_scattering=True):
        self._check(self.lib.mgr_set_scattering(self.h, 1 if use_scattering else 0), "set_scattering")

    def set_window_skip(self, enable=True):
        """Both kernels: eliminate steps out of every line window without
        evaluating their profiles (bit-identical; on by default)."""
        self._check(self.lib.mgr_set_window_skip(self.h, 1 if enable else 0), "set_window_skip")

    def set_trace_options(self, record_steps=-1, candidate_masks=True):
        """Options of the next trace: cells recorded per walk for the filling
        pass (-1 default, 0 none) and the candidate masks (bit-identical
        traces either way)."""
        self._check(self.lib.mgr_set_trace_options(self.h, int(record_steps), 1 if candidate_masks else 0),
                    "set_trace_options")

    def set_far_cache(self, enable=True):
        """FP32 kernel: keep the far-end intensities of every ray and bin
        instead of searching for them in each radiation field (bit-identical)."""
        self._check(self.lib.mgr_set_far_cache(self.h, 1 if enable else 0), "set_far_cache")

    def window_skip_stats(self):
        """(point-chunks of the trace, point-chunks the FP32 kernel skips)."""
        a, b = C.c_int64(0), C.c_int64(0)
        self._check(self.lib.mgr_window_skip_stats(self.h, C.byref(a), C.byref(b)), "window_skip_stats")
        return a.value, b.value

    def u(self):
        """Radiation::u [nrays/2, N, F] of the last radiation field (scattering on)."""
        a = np.zeros((self.R // 2) * self.N * self.F)
        self._check(self.lib.mgr_get_u(self.h, _p(a)), "get_u")
        return a.reshape(self.R // 2, self.N, self.F)

    def npoints_on_rays_rest(self):
        a = np.zeros((self.R // 2) * self.N, np.int64)
        self._check(self.lib.mgr_get_npoints_on_rays_rest(self.h, _p(a)), "get_npoints_on_rays_rest")
        return a.reshape(self.R // 2, self.N)

    def compute_image(self, ray, precision="fp64"):
        """(I_m, I_p), each [ncells, nfreqs], of the image along ray < nrays/2."""
        I_m = np.zeros((self.N, self.F))
        I_p = np.zeros((self.N, self.F))
        self._check(self.lib.mgr_compute_image(self.h, int(ray), _prec(precision), _p(I_m), _p(I_p)),
                    "compute_image")
        return I_m, I_p

    def last_radiation_field_ms(self):
        v = C.c_double(0.0)
        self._check(self.lib.mgr_last_radiation_field_ms(self.h, C.byref(v)), "last_radiation_field_ms")
        return v.value

    def last_radiation_field_kernel(self):
        buf = C.create_string_buffer(128)
        self._check(self.lib.mgr_last_radiation_field_kernel(self.h, buf, 128), "last_radiation_field_kernel")
        return buf.value.decode()

    def radiation_field_shape(self, precision="fp32"):
        """(waves per workgroup, steps per chunk) chosen by the autotuner, and
        the shapes it tried with their times (ms)."""
        w, c, n = C.c_int(0), C.c_int(0), C.c_int(0)
        cap = 32
        tw, tc, tm = np.zeros(cap, np.int32), np.zeros(cap, np.int32), np.zeros(cap, np.float32)
        self._check(self.lib.mgr_radiation_field_shape(self.h, _prec(precision), C.byref(w), C.byref(c), cap,
                                                       C.byref(n), _p(tw), _p(tc), _p(tm)), "radiation_field_shape")
        k = min(n.value, cap)
        return (w.value, c.value), [(int(tw[i]), int(tc[i]), float(tm[i])) for i in range(k)]

    def set_radiation_field_shape(self, precision="fp32", waves_per_group=0, steps_per_chunk=0):
        """Force the launch shape (0, 0: back to the autotuner)."""
        self._check(self.lib.mgr_set_radiation_field_shape(self.h, _prec(precision), int(waves_per_group),
                                                           int(steps_per_chunk)), "set_radiation_field_shape")

    def trace_stats(self):
        a, b = C.c_int64(0), C.c_int64(0)
        self._check(self.lib.mgr_trace_stats(self.h, C.byref(a), C.byref(b)), "trace_stats")
        return a.value, b.value

    # ---- learned warm start ----------------------------------------------------
    def features(self):
        """Per-cell warm-start features [N, 8] (after a radiation field and J_eff)."""
        a = np.zeros((self.N, MLP_IN), np.float32)
        self._check(self.lib.mgr_features(self.h, _p(a)), "features")
        return a

    def prepare(self, precision="fp32"):
        """Per-model caches, tuning and ALI graphs of a precision, built ahead."""
        self._check(self.lib.mgr_prepare(self.h, _prec(precision)), "prepare")

    def first_pass(self, precision="fp32", pair_stride=1):
        """Radiation field over every pair_stride-th ray pair, then J_eff."""
        self._check(self.lib.mgr_first_pass(self.h, _prec(precision), int(pair_stride)), "first_pass")

    def warm_start(self, mlp):
        """Replace the populations by the network's prediction; returns the
        number of cells whose prediction was rejected (not finite and
        positive, or a line with opacity <= 0) and that kept their previous
        populations."""
        self._check(self.lib.mgr_warm_start(self.h, mlp.h), "warm_start")
        n = C.c_int64(0)
        self._check(self.lib.mgr_warm_start_fallbacks(self.h, C.byref(n)), "warm_start_fallbacks")
        return n.value


class BatchContext(Context):
    """Many independent models of the same species in one context
    (mgr_create_batch): concatenated cells, one launch per kernel."""

    def __init__(self, models, device=0):
        from .model import ModelView
        self.lib = load()
        self.models = list(models)
        views, keeps = zip(*[m.view() for m in self.models])
        self._keep = keeps
        arr = (ModelView * len(views))(*views)
        self._views = arr
        h = C.c_void_p()
        rc = self.lib.mgr_create_batch(C.cast(arr, C.c_void_p), len(views), device, C.byref(h))
        if rc != 0:
            raise HipError(f"mgr_create_batch failed ({rc}): {self.lib.mgr_last_error(None).decode()}")
        self.h = h
        m0 = self.models[0]
        self.model = m0
        self.N = sum(m.ncells for m in self.models)
        self.R = max(m.nrays for m in self.models)
        self.F, self.L, self.V, self.Q = m0.nfreqs, m0.nlines, m0.nlev, m0.nquads
        self.Vs = [m0.nlev]
        self.NB = sum(m.nboundary for m in self.models)
        self.M = len(self.models)
        off = np.zeros(self.M + 1, np.int64)
        n = C.c_int64(0)
        self._check(self.lib.mgr_batch_info(self.h, C.byref(n), _p(off)), "batch_info")
        self.cell_offsets = off

    def split(self, a):
        """Per-model views of a concatenated per-cell array."""
        return [a[self.cell_offsets[k]:self.cell_offsets[k + 1]] for k in range(self.M)]

    def convergence(self):
        its = np.zeros(self.M, np.int64)
        conv = (ConvStats * self.M)()
        self._check(self.lib.mgr_batch_convergence(self.h, _p(its), C.cast(conv, C.c_void_p)),
                    "batch_convergence")
        return its, [(c.fraction_not_converged, c.relative_change_mean, c.relative_change_max) for c in conv]


MLP_IN, MLP_HIDDEN, MLP_OUT = 8, 128, 7


def _unpack_params(flat):
    """Packed FP32 parameters -> dict W1 b1 W2 b2 W3 b3 ([in, out] weights)."""
    shapes = [("W1", (MLP_IN, MLP_HIDDEN)), ("b1", (MLP_HIDDEN,)), ("W2", (MLP_HIDDEN, MLP_HIDDEN)),
              ("b2", (MLP_HIDDEN,)), ("W3", (MLP_HIDDEN, MLP_OUT)), ("b3", (MLP_OUT,))]
    out, o = {}, 0
    for k, sh in shapes:
        n = int(np.prod(sh))
        out[k] = flat[o:o + n].reshape(sh)
        o += n
    return out


def _pack_params(p):
    return np.concatenate([np.asarray(p[k], np.float32).ravel() for k in ("W1", "b1", "W2", "b2", "W3", "b3")])


class Mlp:
    """The warm-start network (mgr_mlp) on one device."""

    def __init__(self, seed=0, device=0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.mgr_mlp_create(device, seed, C.byref(h))
        if rc != 0:
            raise HipError(f"mgr_mlp_create failed ({rc})")
        self.h = h
        self.nparams = int(self.lib.mgr_mlp_num_params())

    def close(self):
        if getattr(self, "h", None):
            self.lib.mgr_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            raise HipError(f"{what} failed ({rc}): {self.lib.mgr_mlp_last_error(self.h).decode()}")

    def params(self):
        a = np.zeros(self.nparams, np.float32)
        self._check(self.lib.mgr_mlp_get_params(self.h, _p(a)), "get_params")
        return _unpack_params(a)

    def set_params(self, p):
        a = np.ascontiguousarray(_pack_params(p), dtype=np.float32)
        assert a.size == self.nparams
        self._check(self.lib.mgr_mlp_set_params(self.h, _p(a)), "set_params")

    def set_jlin_inputs(self, use):
        """Whether the network reads the first-pass Jlin inputs (4..7)."""
        self._check(self.lib.mgr_mlp_set_jlin_inputs(self.h, 1 if use else 0), "set_jlin_inputs")

    def set_normalisation(self, x_mean, x_std, y_mean, y_std):
        arrs = [np.ascontiguousarray(v, dtype=np.float32) for v in (x_mean, x_std, y_mean, y_std)]
        self._check(self.lib.mgr_mlp_set_normalisation(self.h, *[_p(v) for v in arrs]), "set_normalisation")

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], MLP_OUT), np.float32)
        self._check(self.lib.mgr_mlp_forward(self.h, _p(X), X.shape[0], _p(Y)), "forward")
        return Y

    def loss_grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.zeros(self.nparams, np.float32)
        loss = C.c_float(0.0)
        self._check(self.lib.mgr_mlp_loss_grad(self.h, _p(X), _p(T), X.shape[0], C.byref(loss), _p(g)),
                    "loss_grad")
        return loss.value, _unpack_params(g)

    def train(self, X, T, epochs, batch, lr, seed=0):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        losses = np.zeros(epochs, np.float32)
        self._check(self.lib.mgr_mlp_train(self.h, _p(X), _p(T), X.shape[0], epochs, batch, lr, seed,
                                           _p(losses)), "train")
        return losses
