"""Fill the drop-in ``core.Simulation`` from an in-memory model.

The reference's Python set-up writes a model to disk and the C++ side reads
it back (src/Model/model.cpp:54-86); ``to_simulation`` fills the same members
directly, so that a model built in Python can go through the drop-in host layer
(``magritte_amd.core``) with or without a file in between.
"""
import numpy as np

from .model import Model


def core():
    """The compiled host-layer module; raises if it was not built."""
    try:
        from . import core as _core
    except ImportError as exc:   # pragma: no cover - message path
        raise ImportError("magritte_amd.core is not built: run python -m magritte_amd.build") from exc
    return _core


def to_simulation(m: Model, use_scattering: bool = False):
    c = core()
    sim = c.Simulation()
    N, R, NB = m.ncells, m.nrays, m.nboundary
    p = sim.parameters
    p.set_ncells(N)
    p.set_nrays(R)
    p.set_nrays_red(R // 2)
    p.set_nboundary(NB)
    p.set_nspecs(len(m.species))
    p.set_nlspecs(m.nlspecs)
    p.set_nlines(m.nlines)
    p.set_nquads(m.nquads)
    p.set_nfreqs(m.nfreqs)
    p.set_nfreqs_red(m.nfreqs)
    p.set_pop_prec(float(m.pop_prec))
    p.max_width_fraction = float(m.max_width_fraction)
    p.n_off_diag = int(m.n_off_diag)
    p.set_use_scattering(bool(use_scattering))

    cells = sim.geometry.cells
    cells.position = np.asarray(m.position, dtype=np.float64).tolist()
    cells.velocity = np.asarray(m.velocity, dtype=np.float64).tolist()
    nn = np.asarray(m.n_neighbors, dtype=np.int64)
    cells.n_neighbors = nn.tolist()
    off = np.concatenate([[0], np.cumsum(nn)])
    nb = np.asarray(m.neighbors, dtype=np.int64)
    cells.neighbors = [nb[off[i]:off[i + 1]].tolist() for i in range(N)]

    rays = sim.geometry.rays
    rays.rays = np.asarray(m.rays, dtype=np.float64).tolist()
    rays.weights = np.asarray(m.weights, dtype=np.float64).tolist()
    rays.nrays = R
    rays.half_nrays = R // 2

    bd = sim.geometry.boundary
    b2c = np.asarray(m.boundary2cell_nr, dtype=np.int64)
    bd.boundary2cell_nr = b2c.tolist()
    c2b = np.full(N, NB, dtype=np.int64)
    c2b[b2c] = np.arange(NB)
    bd.cell2boundary_nr = c2b.tolist()
    flag = np.zeros(N, dtype=bool)
    flag[b2c] = True
    bd.boundary = flag.tolist()
    bd.boundary_condition = [c.BoundaryCondition.CMB] * NB
    tb = m.boundary_temperature if m.boundary_temperature is not None else np.full(NB, 2.72548)
    bd.boundary_temperature = np.asarray(tb, dtype=np.float64).tolist()

    sim.thermodynamics.temperature.gas = np.asarray(m.temperature, dtype=np.float64).tolist()
    sim.thermodynamics.turbulence.vturb2 = np.asarray(m.vturb2, dtype=np.float64).tolist()
    sim.chemistry.species.abundance = np.asarray(m.abundance, dtype=np.float64).tolist()
    sim.chemistry.species.sym = list(m.species)

    species = [(m.linedata, m.roots, m.qweights)] + [(x.linedata, x.roots, x.qweights) for x in m.more_species]
    for ld, roots, qweights in species:
        sim.lines.lineProducingSpecies.append(_line_species(c, m, ld, roots, qweights))
    freqs = np.concatenate([ld.frequency for ld, _, _ in species])
    sim.lines.line = sorted(freqs.tolist())
    sim.lines.line_index = [int(i) for i in np.argsort(freqs, kind="stable")]
    if use_scattering:   # Radiation::read sizes u (radiation.cpp:117-126)
        sim.radiation.u = [[0.0] * (N * m.nfreqs) for _ in range(R // 2)]
    return sim


def _line_species(c, m, ld, roots, qweights):
    """One Lines/LineProducingSpecies_<l> of the host Model, sized as
    LineProducingSpecies::read leaves it (lineProducingSpecies.cpp:21-70)."""
    N = m.ncells
    lsp = c.LineProducingSpecies()
    cl = c.Linedata()
    cl.num = int(ld.num)
    cl.sym = ld.sym
    cl.inverse_mass = float(ld.inverse_mass)
    cl.nlev = ld.nlev
    cl.nrad = ld.nrad
    cl.irad = np.asarray(ld.irad, dtype=np.int64).tolist()
    cl.jrad = np.asarray(ld.jrad, dtype=np.int64).tolist()
    cl.energy = np.asarray(ld.energy, dtype=np.float64).tolist()
    cl.weight = np.asarray(ld.weight, dtype=np.float64).tolist()
    cl.frequency = ld.frequency.tolist()
    cl.A = np.asarray(ld.A, dtype=np.float64).tolist()
    cl.Ba = ld.Ba.tolist()
    cl.Bs = ld.Bs.tolist()
    cl.ncolpar = len(ld.colpar)
    cps = c.vCollisionPartner()
    for cp in ld.colpar:
        q = c.CollisionPartner()
        q.num_col_partner = int(cp.num)
        q.orth_or_para_H2 = cp.orth_or_para
        q.ntmp = int(len(cp.tmp))
        q.ncol = int(len(cp.icol))
        q.icol = np.asarray(cp.icol, dtype=np.int64).tolist()
        q.jcol = np.asarray(cp.jcol, dtype=np.int64).tolist()
        q.tmp = np.asarray(cp.tmp, dtype=np.float64).tolist()
        q.Cd = np.asarray(cp.Cd, dtype=np.float64).tolist()
        q.Ce = np.asarray(cp.Ce, dtype=np.float64).tolist()
        cps.append(q)
    cl.colpar = cps
    lsp.linedata = cl
    qd = c.Quadrature()
    qd.roots = np.asarray(roots, dtype=np.float64).tolist()
    qd.weights = np.asarray(qweights, dtype=np.float64).tolist()
    lsp.quadrature = qd
    lsp.ncells = N
    zeros_nv = [0.0] * (N * ld.nlev)
    lsp.population = zeros_nv
    lsp.population_prev1 = zeros_nv
    lsp.population_prev2 = zeros_nv
    lsp.population_prev3 = zeros_nv
    lsp.population_tot = np.asarray(m.abundance, dtype=np.float64)[:, ld.num].tolist()
    lsp.Jlin = [[0.0] * ld.nrad for _ in range(N)]
    lsp.Jeff = [[0.0] * ld.nrad for _ in range(N)]
    lsp.Jdif = [[0.0] * ld.nrad for _ in range(N)]
    return lsp
