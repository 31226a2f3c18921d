"""ctypes binding of oracle/liboracle.so (the CPU FP64 restatement).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

How each oracle function is pinned (tests/test_oracle_*.py):
  orc_raypair_solve          -- u = T^-1 S' and L = diag/bands of T^-1 for the
                                Feautrier matrix T (docs/background.tex:303-423),
                                dense numpy inverse, incl. the known-answer set-up
                                of tests/unit/Simulation/Raypair/test_raypair.cpp;
                                constant-source slab closed form.
  orc_get_doppler_shift      -- known answers of tests/unit/Model/Geometry/test_geometry.cpp
  orc_trace_ray              -- invariants (Z increasing, neighbours, ends on the
                                boundary, point counts == orc_get_npoints_on_ray),
                                axis-aligned rays on a regular grid (brute force)
                                and an independent numpy restatement.
  orc_get_eta_and_chi        -- numpy brute force over all lines incl. the window
                                and last-line rule (P2) and the 1e-26 floor (P3).
  spectral discretisation, boundary intensities, LTE -- closed forms (numpy).
  Jeff, stat-eq              -- closed-form sums; 2-level analytic populations;
                                detailed-balance invariant of
                                tests/unit/Model/Lines/test_lineProducingSpecies.cpp.
  Ng                         -- numpy restatement of the Ng formula and the
                                exact-extrapolation property on geometric sequences.
  collision rates            -- closed-form linear interpolation/extrapolation.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(path)
        vp = C.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [vp]
        L.orc_destroy.argtypes = [vp]
        for name in ("orc_compute_spectral_discretisation", "orc_compute_LTE_level_populations",
                     "orc_update_using_statistical_equilibrium"):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = C.c_int
        L.orc_compute_boundary_intensities.argtypes = [vp, vp]
        L.orc_compute_boundary_intensities.restype = C.c_int
        L.orc_set_emissivity_and_opacity.argtypes = [vp]
        L.orc_rescale_I_bdy.argtypes = [vp, C.c_double, C.c_int64, C.c_int64, C.c_int64]
        L.orc_rescale_I_bdy.restype = C.c_double
        L.orc_cell2boundary_nr.argtypes = [vp, C.c_int64]
        L.orc_cell2boundary_nr.restype = C.c_int64
        L.orc_get_dshift_max.argtypes = [vp, C.c_int64]
        L.orc_get_dshift_max.restype = C.c_double
        L.orc_trace_ray.argtypes = [vp, C.c_int64, C.c_int64, C.c_double, C.c_int64, vp, vp, vp, vp]
        L.orc_trace_ray.restype = C.c_int64
        L.orc_trace_ray_rest.argtypes = [vp, C.c_int64, C.c_int64, C.c_double, C.c_int64, vp, vp, vp, vp]
        L.orc_trace_ray_rest.restype = C.c_int64
        L.orc_compute_spectral_discretisation_image.argtypes = [vp, C.c_double]
        L.orc_compute_spectral_discretisation_image.restype = C.c_int
        L.orc_solve_for_image.argtypes = [C.c_int64, C.c_int64, C.c_int64, vp, vp, C.c_double, C.c_double, vp]
        L.orc_compute_image.argtypes = [vp, C.c_int64, vp, vp]
        L.orc_compute_image.restype = C.c_int
        L.orc_get_npoints_on_ray.argtypes = [vp, C.c_int64, C.c_int64, C.c_double]
        L.orc_get_npoints_on_ray.restype = C.c_int64
        L.orc_get_doppler_shift.argtypes = [vp, C.c_int64, C.c_int64, C.c_int64, C.c_int]
        L.orc_get_doppler_shift.restype = C.c_double
        L.orc_raypair_solve.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int64,
                                        vp, vp, C.c_double, C.c_double, vp, vp, vp, vp, vp, vp]
        L.orc_compute_radiation_field.argtypes = [vp, C.c_int, C.c_int64, C.c_int64]
        L.orc_compute_radiation_field.restype = C.c_int
        L.orc_compute_radiation_field_list.argtypes = [vp, C.c_int, vp, C.c_int64]
        L.orc_compute_radiation_field_list.restype = C.c_int
        L.orc_compute_Jeff.argtypes = [vp]
        L.orc_check_for_convergence.argtypes = [vp, C.c_double]
        L.orc_update_using_Ng_acceleration.argtypes = [vp]
        L.orc_compute_level_populations.argtypes = [vp, C.c_int, C.c_int64, C.c_int, C.c_int]
        L.orc_compute_level_populations.restype = C.c_int64
        L.orc_nthreads_max.restype = C.c_int
        for name in ("orc_get_J", "orc_get_lambda", "orc_get_populations", "orc_set_populations",
                     "orc_get_boundary_intensities", "orc_get_antipod"):
            getattr(L, name).argtypes = [vp, vp]
        L.orc_get_frequencies.argtypes = [vp, vp, vp]
        L.orc_get_Jeff.argtypes = [vp, vp, vp, vp]
        L.orc_lambda_band_capacity.argtypes = [vp]
        L.orc_lambda_band_capacity.restype = C.c_int64
        L.orc_get_lambda_band.argtypes = [vp, vp, vp, vp]
        L.orc_set_scattering.argtypes = [vp, C.c_int]
        L.orc_get_u.argtypes = [vp, vp]
        L.orc_get_population_history.argtypes = [vp, vp, vp, vp]
        L.orc_set_population_history.argtypes = [vp, vp, vp, vp]
        L.orc_get_emissivity_opacity.argtypes = [vp, vp, vp]
        L.orc_get_convergence.argtypes = [vp, vp, vp, vp]
        L.orc_nlspecs.argtypes = [vp]
        L.orc_nlspecs.restype = C.c_int64
        L.orc_get_convergence_species.argtypes = [vp, C.c_int64, vp, vp, vp]
        L.orc_get_lines.argtypes = [vp, vp, vp]
        L.orc_get_line_maps.argtypes = [vp, vp, vp, vp, vp]
        L.orc_planck.argtypes = [C.c_double, C.c_double]
        L.orc_planck.restype = C.c_double
        L.orc_profile.argtypes = [C.c_double, C.c_double]
        L.orc_profile.restype = C.c_double
        L.orc_profile_width.argtypes = [vp, C.c_int64]
        L.orc_profile_width.restype = C.c_double
        L.orc_get_eta_and_chi.argtypes = [vp, C.c_double, C.c_int64, vp, vp, vp]
        L.orc_collision_rates.argtypes = [vp, C.c_int64, C.c_double, vp, vp, vp]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """FP64 CPU restatement bound to one model."""

    def __init__(self, model):
        self.model = model
        self._view, self._keep = model.view()
        self._s = lib().orc_create(C.byref(self._view))
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

    def __del__(self):
        if getattr(self, "_s", None):
            lib().orc_destroy(self._s)
            self._s = None

    # set-up
    def compute_spectral_discretisation(self):
        return lib().orc_compute_spectral_discretisation(self._s)

    def compute_boundary_intensities(self, temperatures=None):
        t = None if temperatures is None else np.ascontiguousarray(temperatures, dtype=np.float64)
        return lib().orc_compute_boundary_intensities(self._s, None if t is None else _p(t))

    def compute_LTE_level_populations(self):
        return lib().orc_compute_LTE_level_populations(self._s)

    def setup(self):
        self.compute_spectral_discretisation()
        self.compute_boundary_intensities()
        self.compute_LTE_level_populations()
        return self

    def set_emissivity_and_opacity(self):
        lib().orc_set_emissivity_and_opacity(self._s)

    def rescale_I_bdy(self, freq_scaled, R, p):
        """Boundary intensity of boundary cell p for reduced ray R at freq_scaled."""
        return lib().orc_rescale_I_bdy(self._s, float(freq_scaled), int(R), int(p),
                                       lib().orc_cell2boundary_nr(self._s, int(p)))

    def cell2boundary_nr(self, p):
        return lib().orc_cell2boundary_nr(self._s, int(p))

    # geometry
    def dshift_max(self, o):
        return lib().orc_get_dshift_max(self._s, o)

    def compute_spectral_discretisation_image(self, width):
        return lib().orc_compute_spectral_discretisation_image(self._s, width)

    def compute_image(self, r):
        """(I_m, I_p) [N, F] of the image along ray r < nrays/2."""
        Im = np.zeros(self.N * self.F); Ip = np.zeros(self.N * self.F)
        if lib().orc_compute_image(self._s, r, _p(Im), _p(Ip)) != 0:
            raise RuntimeError("oracle image failed (trace error or ray out of range)")
        return Im.reshape(self.N, self.F), Ip.reshape(self.N, self.F)

    def trace_ray(self, o, r, dshift_max=None, rest=False):
        if dshift_max is None:
            dshift_max = self.dshift_max(o)
        fn = lib().orc_trace_ray_rest if rest else lib().orc_trace_ray
        cap = 64
        while True:
            crt = np.zeros(cap, np.int64); cell = np.zeros(cap, np.int64)
            sh = np.zeros(cap); dz = np.zeros(cap)
            n = fn(self._s, o, r, dshift_max, cap, _p(crt), _p(cell), _p(sh), _p(dz))
            if n < 0:
                raise RuntimeError("ray found no forward neighbour")
            if n <= cap:
                return crt[:n], cell[:n], sh[:n], dz[:n]
            cap = 2 * n

    def npoints_on_ray(self, o, r, dshift_max=None):
        if dshift_max is None:
            dshift_max = self.dshift_max(o)
        return lib().orc_get_npoints_on_ray(self._s, o, r, dshift_max)

    def doppler_shift(self, o, r, c, rest=False):
        return lib().orc_get_doppler_shift(self._s, o, r, c, 1 if rest else 0)

    def antipod(self):
        a = np.zeros(self.R, np.int64)
        lib().orc_get_antipod(self._s, _p(a))
        return a

    # hot path
    def compute_radiation_field(self, nthreads=0, o_begin=0, o_end=-1):
        rc = lib().orc_compute_radiation_field(self._s, nthreads, o_begin, o_end)
        if rc != 0:
            raise RuntimeError("oracle radiation field failed (trace error)")

    def compute_radiation_field_list(self, origins, nthreads=0):
        o = np.ascontiguousarray(origins, dtype=np.int64)
        rc = lib().orc_compute_radiation_field_list(self._s, nthreads, _p(o), o.shape[0])
        if rc != 0:
            raise RuntimeError("oracle radiation field failed (trace error)")

    def compute_Jeff(self):
        lib().orc_compute_Jeff(self._s)

    def update_using_statistical_equilibrium(self):
        if lib().orc_update_using_statistical_equilibrium(self._s) != 0:
            raise RuntimeError("singular rate matrix")

    def check_for_convergence(self, pop_prec=None):
        lib().orc_check_for_convergence(self._s, self.model.pop_prec if pop_prec is None else pop_prec)
        return self.convergence()

    def update_using_Ng_acceleration(self):
        lib().orc_update_using_Ng_acceleration(self._s)

    def compute_level_populations(self, use_Ng=True, max_niterations=1000, criterion=0, nthreads=0):
        return lib().orc_compute_level_populations(self._s, 1 if use_Ng else 0, max_niterations,
                                                   criterion, nthreads)

    # accessors
    def frequencies(self):
        nu = np.zeros(self.N * self.F); nr = np.zeros(self.N * self.L * self.Q, np.int64)
        lib().orc_get_frequencies(self._s, _p(nu), _p(nr))
        return nu.reshape(self.N, self.F), nr.reshape(self.N, self.L, self.Q)

    def boundary_intensities(self):
        a = np.zeros(self.NB * self.F)
        lib().orc_get_boundary_intensities(self._s, _p(a))
        return a.reshape(self.NB, self.F)

    def set_scattering(self, use_scattering=True):
        lib().orc_set_scattering(self._s, 1 if use_scattering else 0)

    def u(self):
        """Radiation::u [nrays/2, N, F] of the last radiation field (scattering on)."""
        a = np.zeros((self.R // 2) * self.N * self.F)
        lib().orc_get_u(self._s, _p(a))
        return a.reshape(self.R // 2, self.N, self.F)

    def J(self):
        a = np.zeros(self.N * self.F)
        lib().orc_get_J(self._s, _p(a))
        return a.reshape(self.N, self.F)

    def lambda_diag(self):
        a = np.zeros(self.N * self.L)
        lib().orc_get_lambda(self._s, _p(a))
        return a.reshape(self.N, self.L)

    def lambda_band(self):
        """Lambda* elements with nr != p (n_off_diag >= 2): a list per (p, k) of
        (nr, Ls) in Lambda::add_element order."""
        cap = lib().orc_lambda_band_capacity(self._s)
        out = [[[] for _ in range(self.L)] for _ in range(self.N)]
        if cap == 0:
            return out
        cnt = np.zeros(self.N * self.L, dtype=np.int64)
        nr = np.zeros(self.N * self.L * cap, dtype=np.int64)
        Ls = np.zeros(self.N * self.L * cap)
        lib().orc_get_lambda_band(self._s, _p(cnt), _p(nr), _p(Ls))
        for p in range(self.N):
            for k in range(self.L):
                pk = p * self.L + k
                out[p][k] = [(int(nr[pk * cap + e]), float(Ls[pk * cap + e])) for e in range(cnt[pk])]
        return out

    def Jeff(self):
        a, b, c = (np.zeros(self.N * self.L) for _ in range(3))
        lib().orc_get_Jeff(self._s, _p(a), _p(b), _p(c))
        return a.reshape(self.N, self.L), b.reshape(self.N, self.L), c.reshape(self.N, self.L)

    def populations(self):
        a = np.zeros(self.N * sum(self.Vs))
        lib().orc_get_populations(self._s, _p(a))
        return self._split(a)

    def set_populations(self, pops):
        if isinstance(pops, (list, tuple)):
            pops = np.concatenate([np.ravel(x) for x in pops])
        a = np.ascontiguousarray(pops, dtype=np.float64).ravel()
        lib().orc_set_populations(self._s, _p(a))

    def population_history(self):
        a, b, c = (np.zeros(self.N * sum(self.Vs)) for _ in range(3))
        lib().orc_get_population_history(self._s, _p(a), _p(b), _p(c))
        return tuple(self._split(x) for x in (a, b, c))

    def convergence_species(self):
        """(fraction not converged, mean change, max change) of each species."""
        out = []
        for l in range(lib().orc_nlspecs(self._s)):
            a, b, c = C.c_double(), C.c_double(), C.c_double()
            lib().orc_get_convergence_species(self._s, l, C.byref(a), C.byref(b), C.byref(c))
            out.append((a.value, b.value, c.value))
        return out

    def set_population_history(self, p1, p2, p3):
        p1, p2, p3 = (np.concatenate([np.ravel(y) for y in x]) if isinstance(x, (list, tuple)) else x
                      for x in (p1, p2, p3))
        arrs = [np.ascontiguousarray(x, dtype=np.float64).ravel() for x in (p1, p2, p3)]
        lib().orc_set_population_history(self._s, *[_p(x) for x in arrs])

    def emissivity_opacity(self):
        a, b = np.zeros(self.N * self.L), np.zeros(self.N * self.L)
        lib().orc_get_emissivity_opacity(self._s, _p(a), _p(b))
        return a.reshape(self.N, self.L), b.reshape(self.N, self.L)

    def convergence(self):
        v = [C.c_double() for _ in range(3)]
        lib().orc_get_convergence(self._s, *[C.byref(x) for x in v])
        return tuple(x.value for x in v)

    def lines(self):
        line = np.zeros(self.L); idx = np.zeros(self.L, np.int64)
        lib().orc_get_lines(self._s, _p(line), _p(idx))
        return line, idx

    def line_maps(self):
        arrs = [np.zeros(self.F, np.int64) for _ in range(4)]
        lib().orc_get_line_maps(self._s, *[_p(a) for a in arrs])
        return arrs

    def eta_and_chi(self, freq_scaled, p, lnotch=0):
        ln = C.c_int64(lnotch); eta = C.c_double(); chi = C.c_double()
        lib().orc_get_eta_and_chi(self._s, freq_scaled, p, C.byref(ln), C.byref(eta), C.byref(chi))
        return eta.value, chi.value, ln.value

    def profile_width(self, p):
        return lib().orc_profile_width(self._s, p)

    def collision_rates(self, c, temperature):
        cp = self.model.linedata.colpar[c]
        n = len(cp.icol)
        ce, cd = np.zeros(n), np.zeros(n)
        fac = C.c_double()
        lib().orc_collision_rates(self._s, c, temperature, _p(ce), _p(cd), C.byref(fac))
        return ce, cd, fac.value


def raypair_solve(term1, dtau, I0, In, n_ar, first=None, last=None, n_off_diag=0):
    """RayPair::solve on given inputs (raypair_solver.tpp:200-420)."""
    term1 = np.ascontiguousarray(term1, dtype=np.float64)
    dtau = np.ascontiguousarray(dtau, dtype=np.float64)
    ndep = term1.shape[0]
    first = 0 if first is None else first
    last = ndep - 1 if last is None else last
    Su = np.zeros(ndep); Ld = np.zeros(ndep)
    Lu = np.zeros(max(n_off_diag, 1) * ndep); Ll = np.zeros(max(n_off_diag, 1) * ndep)
    A = np.zeros(ndep); Cc = np.zeros(ndep)
    lib().orc_raypair_solve(ndep, n_ar, first, last, n_off_diag, _p(term1), _p(dtau), I0, In,
                            _p(Su), _p(Ld), _p(Lu), _p(Ll), _p(A), _p(Cc))
    return dict(Su=Su, L_diag=Ld, L_upper=Lu.reshape(-1, ndep), L_lower=Ll.reshape(-1, ndep),
                A=A, C=Cc)


def solve_for_image(term1, dtau, I0, In, first=None, last=None):
    """RayPair::solve_for_image on given inputs (raypair_solver.tpp:523-579): u at every point."""
    term1 = np.ascontiguousarray(term1, dtype=np.float64)
    dtau = np.ascontiguousarray(dtau, dtype=np.float64)
    ndep = term1.shape[0]
    first = 0 if first is None else first
    last = ndep - 1 if last is None else last
    Su = np.zeros(ndep)
    lib().orc_solve_for_image(ndep, first, last, _p(term1), _p(dtau), I0, In, _p(Su))
    return Su


def planck(T, nu):
    return lib().orc_planck(T, nu)


def profile(width, freq_diff):
    return lib().orc_profile(width, freq_diff)
