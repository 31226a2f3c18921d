"""In-memory model in the reference io layout, and its C-ABI view.

Field names follow the reference's io model layout (Model::read,
src/Model/model.cpp:54-86; group/dataset names in the respective read()
functions).  ``ModelView`` mirrors ``mgr_model_view`` of include/magritte_hip.h
field for field.
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .setup.linedata import Linedata

T_CMB = 2.7254800   # src/Tools/constants.hpp:38

_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)


_SPECIES_FIELDS = [
    ("num", C.c_int64), ("inverse_mass", C.c_double), ("nlev", C.c_int64), ("nrad", C.c_int64),
    ("energy", _f64p), ("weight", _f64p), ("irad", _i64p), ("jrad", _i64p),
    ("A", _f64p), ("Bs", _f64p), ("Ba", _f64p), ("frequency", _f64p),
    ("ncolpar", C.c_int64), ("colpar_num", _i64p), ("colpar_orth_or_para", C.c_char_p),
    ("colpar_ncol", _i64p), ("colpar_ntmp", _i64p), ("icol", _i64p), ("jcol", _i64p),
    ("tmp", _f64p), ("Cd", _f64p), ("Ce", _f64p),
    ("roots", _f64p), ("qweights", _f64p),
]


class SpeciesView(C.Structure):
    _fields_ = _SPECIES_FIELDS


class ModelView(C.Structure):
    _fields_ = [
        ("ncells", C.c_int64), ("nrays", C.c_int64), ("nboundary", C.c_int64),
        ("nspecs", C.c_int64), ("nquads", C.c_int64),
        ("pop_prec", C.c_double), ("max_width_fraction", C.c_double),
        ("n_off_diag", C.c_int64),
        ("position", _f64p), ("velocity", _f64p), ("n_neighbors", _i64p), ("neighbors", _i64p),
        ("rays", _f64p), ("weights", _f64p),
        ("boundary2cell_nr", _i64p),
        ("temperature", _f64p), ("vturb2", _f64p),
        ("abundance", _f64p),
        ("num", C.c_int64), ("inverse_mass", C.c_double), ("nlev", C.c_int64), ("nrad", C.c_int64),
        ("energy", _f64p), ("weight", _f64p), ("irad", _i64p), ("jrad", _i64p),
        ("A", _f64p), ("Bs", _f64p), ("Ba", _f64p), ("frequency", _f64p),
        ("ncolpar", C.c_int64), ("colpar_num", _i64p), ("colpar_orth_or_para", C.c_char_p),
        ("colpar_ncol", _i64p), ("colpar_ntmp", _i64p), ("icol", _i64p), ("jcol", _i64p),
        ("tmp", _f64p), ("Cd", _f64p), ("Ce", _f64p),
        ("roots", _f64p), ("qweights", _f64p),
        ("nlspecs", C.c_int64), ("more_species", C.POINTER(SpeciesView)),
    ]


@dataclass
class LineSpecies:
    """A further line-producing species (Lines/LineProducingSpecies_<l>, l >= 1):
    its line data and quadrature (nquads roots and weights, as species 0's)."""
    linedata: Linedata
    roots: np.ndarray
    qweights: np.ndarray


class ConvStats(C.Structure):
    _fields_ = [("fraction_not_converged", C.c_double),
                ("relative_change_mean", C.c_double),
                ("relative_change_max", C.c_double)]


@dataclass
class Model:
    """Model arrays in the reference io layout (SI units, velocities in c)."""
    position: np.ndarray             # Geometry/Cells/position [N,3]
    velocity: np.ndarray             # Geometry/Cells/velocity [N,3]
    n_neighbors: np.ndarray          # Geometry/Cells/n_neighbors [N]
    neighbors: np.ndarray            # Geometry/Cells/neighbors (flat)
    rays: np.ndarray                 # Geometry/Rays/rays [R,3]
    weights: np.ndarray              # Geometry/Rays/weights [R]
    boundary2cell_nr: np.ndarray     # Geometry/Boundary/boundary2cell_nr
    temperature: np.ndarray          # Thermodynamics/Temperature/gas
    vturb2: np.ndarray               # Thermodynamics/Turbulence/vturb2
    abundance: np.ndarray            # Chemistry/Species/abundance [N, nspecs]
    species: List[str]
    linedata: Linedata
    roots: np.ndarray                # Lines/LineProducingSpecies_0/Quadrature/roots
    qweights: np.ndarray             # .../Quadrature/weights
    boundary_temperature: Optional[np.ndarray] = None
    pop_prec: float = 1.0e-6
    max_width_fraction: float = 0.5
    n_off_diag: int = 0
    meta: dict = field(default_factory=dict)
    more_species: List[LineSpecies] = field(default_factory=list)

    @property
    def ncells(self): return int(self.position.shape[0])
    @property
    def nrays(self): return int(self.rays.shape[0])
    @property
    def nboundary(self): return int(self.boundary2cell_nr.shape[0])
    @property
    def nquads(self): return int(self.roots.shape[0])
    @property
    def nlev(self): return self.linedata.nlev
    @property
    def nrad(self): return self.linedata.nrad
    @property
    def nfreqs(self): return self.nlines * self.nquads
    @property
    def nlspecs(self): return 1 + len(self.more_species)
    @property
    def nlines(self): return self.nrad + sum(ls.linedata.nrad for ls in self.more_species)
    @property
    def species_linedata(self): return [self.linedata] + [ls.linedata for ls in self.more_species]

    def view(self):
        """Build the C view; the returned keep-alive list must outlive its use."""
        keep = []

        def f64(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(a)
            return a.ctypes.data_as(_f64p)

        def i64(a):
            a = np.ascontiguousarray(a, dtype=np.int64)
            keep.append(a)
            return a.ctypes.data_as(_i64p)

        def fill(v, ld, roots, qweights):
            cps = ld.colpar
            op = b"".join(c.orth_or_para.encode()[:1] for c in cps) + b"\0"
            keep.append(op)
            v.num, v.inverse_mass = ld.num, ld.inverse_mass
            v.nlev, v.nrad = ld.nlev, ld.nrad
            v.energy, v.weight = f64(ld.energy), f64(ld.weight)
            v.irad, v.jrad = i64(ld.irad), i64(ld.jrad)
            v.A, v.Bs, v.Ba, v.frequency = f64(ld.A), f64(ld.Bs), f64(ld.Ba), f64(ld.frequency)
            v.ncolpar = len(cps)
            v.colpar_num = i64(np.array([c.num for c in cps] or [0]))
            v.colpar_orth_or_para = C.c_char_p(op)
            v.colpar_ncol = i64(np.array([len(c.icol) for c in cps] or [0]))
            v.colpar_ntmp = i64(np.array([len(c.tmp) for c in cps] or [0]))
            v.icol = i64(np.concatenate([c.icol for c in cps]) if cps else np.zeros(1))
            v.jcol = i64(np.concatenate([c.jcol for c in cps]) if cps else np.zeros(1))
            v.tmp = f64(np.concatenate([c.tmp for c in cps]) if cps else np.zeros(1))
            v.Cd = f64(np.concatenate([c.Cd.ravel() for c in cps]) if cps else np.zeros(1))
            v.Ce = f64(np.concatenate([c.Ce.ravel() for c in cps]) if cps else np.zeros(1))
            v.roots, v.qweights = f64(roots), f64(qweights)

        v = ModelView()
        v.ncells, v.nrays, v.nboundary = self.ncells, self.nrays, self.nboundary
        v.nspecs = int(self.abundance.shape[1])
        v.nquads = self.nquads
        v.pop_prec, v.max_width_fraction = self.pop_prec, self.max_width_fraction
        v.n_off_diag = self.n_off_diag
        v.position, v.velocity = f64(self.position), f64(self.velocity)
        v.n_neighbors, v.neighbors = i64(self.n_neighbors), i64(self.neighbors)
        v.rays, v.weights = f64(self.rays), f64(self.weights)
        v.boundary2cell_nr = i64(self.boundary2cell_nr)
        v.temperature, v.vturb2 = f64(self.temperature), f64(self.vturb2)
        v.abundance = f64(self.abundance)
        fill(v, self.linedata, self.roots, self.qweights)
        v.nlspecs = 1 + len(self.more_species)
        if self.more_species:
            arr = (SpeciesView * len(self.more_species))()
            for sv, ls in zip(arr, self.more_species):
                fill(sv, ls.linedata, ls.roots, ls.qweights)
            keep.append(arr)
            v.more_species = C.cast(arr, C.POINTER(SpeciesView))
        return v, keep
