"""Seeded synthetic models in the reference io layout.

regular_grid_model   -- the CPU reference configuration: 16^3 regular grid,
                        6-neighbour lists, 2-level line at 115.27 GHz.
envelope_model       -- the at-scale configuration: jittered n^3 point cloud,
                        26 grid-neighbour lists (Voronoi-approximate), spherical
                        envelope, 7-level linear rotor.  n = 64 is the full size;
                        smaller n gives the same physics for parity tests.
stress_model         -- the precision-stress variant of envelope_model: n0 raised
                        to 1e17 m^-3 and 20 % of cells with +-30 km/s velocity jumps.

Fixed choices (not given by the workload description) are recorded in
``Model.meta``: neighbour order, box size, r0 = R* and the seed.
"""
import numpy as np

from ..model import LineSpecies, Model
from . import healpix
from .linedata import CC, linear_rotor, split_rotor, two_level
from .quadrature import gaussian_linspace

# Neighbour offsets of the 6-neighbour lists, in list order (ties in the ray
# tracer go to the first neighbour in this order).
SIX_NEIGHBOURS = ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1))
# 26 grid neighbours: dz outer, dy, dx inner, each over (-1, 0, 1), centre skipped.
TWENTYSIX_NEIGHBOURS = tuple((dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1)
                             for dx in (-1, 0, 1) if (dx, dy, dz) != (0, 0, 0))


def grid_index(n):
    """Cell p = i + n j + n^2 k, x fastest; returns i, j, k of every cell."""
    p = np.arange(n ** 3)
    return p % n, (p // n) % n, p // (n * n)


def grid_neighbors(n, offsets):
    i, j, k = grid_index(n)
    lists = []
    for d in offsets:
        ii, jj, kk = i + d[0], j + d[1], k + d[2]
        ok = (ii >= 0) & (ii < n) & (jj >= 0) & (jj < n) & (kk >= 0) & (kk < n)
        lists.append(np.where(ok, ii + n * jj + n * n * kk, -1))
    nb = np.stack(lists, axis=1)                       # [N, len(offsets)]
    n_neighbors = (nb >= 0).sum(axis=1).astype(np.int64)
    neighbors = nb[nb >= 0].astype(np.int64)           # row-major keeps list order
    return n_neighbors, neighbors


def grid_boundary(n):
    i, j, k = grid_index(n)
    on = (i == 0) | (i == n - 1) | (j == 0) | (j == n - 1) | (k == 0) | (k == n - 1)
    return np.nonzero(on)[0].astype(np.int64)


def regular_grid_model(n=16, side=1.0e14, nside=1, nquads=39, n_H2=1.0e12,
                       abundance=1.0e-4, T=20.0, vturb=150.0, pop_prec=1.0e-6):
    h = side / n
    i, j, k = grid_index(n)
    position = np.stack([(i + 0.5) * h, (j + 0.5) * h, (k + 0.5) * h], axis=1).astype(np.float64)
    N = n ** 3
    n_nb, nbs = grid_neighbors(n, SIX_NEIGHBOURS)
    rays, weights = healpix.ring_rays(nside)
    roots, qw = gaussian_linspace(nquads, 5.0)
    ab = np.zeros((N, 3))
    ab[:, 1] = abundance * n_H2
    ab[:, 2] = n_H2
    return Model(position=position, velocity=np.zeros((N, 3)), n_neighbors=n_nb, neighbors=nbs,
                 rays=rays, weights=weights, boundary2cell_nr=grid_boundary(n),
                 temperature=np.full(N, T), vturb2=np.full(N, (vturb / CC) ** 2),
                 abundance=ab, species=["dummy", "CO", "H2"], linedata=two_level(),
                 roots=roots, qweights=qw, pop_prec=pop_prec,
                 meta=dict(kind="regular_grid", n=n, side=side, neighbours="-x,+x,-y,+y,-z,+z"))


def envelope_model(n=64, side=1.0e15, nside=2, nquads=31, n0=1.0e14, T0=2500.0,
                   v_inf=15.0e3, abundance=1.0e-4, vturb=150.0, jitter=0.3, seed=1234,
                   r0_cells=2.0, jump_fraction=0.0, jump_speed=30.0e3, nlev=7,
                   pop_prec=1.0e-6, linedata=None):
    rng = np.random.default_rng(seed)
    h = side / n
    i, j, k = grid_index(n)
    N = n ** 3
    base = np.stack([(i + 0.5) * h, (j + 0.5) * h, (k + 0.5) * h], axis=1)
    position = (base + rng.uniform(-jitter, jitter, size=(N, 3)) * h).astype(np.float64)
    n_nb, nbs = grid_neighbors(n, TWENTYSIX_NEIGHBOURS)
    centre = 0.5 * side
    d = position - centre
    r = np.sqrt((d * d).sum(axis=1))
    r0 = r0_cells * h
    r_eff = np.maximum(r, r0)
    n_H2 = n0 * (r0 / r_eff) ** 2
    T = T0 * (r0 / r_eff) ** 0.6
    speed = v_inf * np.sqrt(1.0 - r0 / r_eff)                    # m/s, R* = r0
    rhat = np.where(r[:, None] > 0.0, d / np.maximum(r, 1e-300)[:, None], 0.0)
    velocity = (speed[:, None] * rhat) / CC
    if jump_fraction > 0.0:
        jumped = rng.random(N) < jump_fraction
        sign = np.where(rng.random(N) < 0.5, -1.0, 1.0)
        velocity = velocity + (jumped * sign * jump_speed / CC)[:, None] * rhat
    rays, weights = healpix.ring_rays(nside)
    roots, qw = gaussian_linspace(nquads, 5.0)
    ab = np.zeros((N, 3))
    ab[:, 1] = abundance * n_H2
    ab[:, 2] = n_H2
    return Model(position=position, velocity=velocity.astype(np.float64), n_neighbors=n_nb,
                 neighbors=nbs, rays=rays, weights=weights, boundary2cell_nr=grid_boundary(n),
                 temperature=T.astype(np.float64), vturb2=np.full(N, (vturb / CC) ** 2),
                 abundance=ab, species=["dummy", "CO", "H2"], linedata=linear_rotor(nlev=nlev) if linedata is None else linedata,
                 roots=roots, qweights=qw, pop_prec=pop_prec,
                 meta=dict(kind="envelope", n=n, side=side, r0=r0, R_star=r0, n0=n0, T0=T0,
                           seed=seed, jitter=jitter, jump_fraction=jump_fraction,
                           neighbours="26 grid neighbours, dz outer, dy, dx inner"))


def blended_model(n=8, seed=21, **kw):
    """An envelope whose species has blended lines (split_rotor: 10 lines in
    groups of 2 and 4, about one width apart), so that a bin sees up to four
    lines in its window."""
    return envelope_model(n=n, seed=seed, linedata=split_rotor(), **kw)


def stress_model(n=64, seed=4321, **kw):
    kw.setdefault("n0", 1.0e17)
    return envelope_model(n=n, seed=seed, jump_fraction=0.2, **kw)


def two_species_model(n=6, seed=5, second_nu=700.0e9, second_mass=28.0, second_abundance=3.0e-5,
                      second_A=2.0e-6, **kw):
    """The envelope with a 4-level rotor (lines at 115, 230 and 346 GHz) and a
    second line-producing species: a two-level species at second_nu with its
    own abundance (species index 3, collision partner H2), so
    Lines/LineProducingSpecies_0 and _1."""
    kw.setdefault("nlev", 4)
    m = envelope_model(n=n, seed=seed, **kw)
    extra = (second_abundance * m.abundance[:, 2])[:, None]
    m.abundance = np.concatenate([m.abundance, extra], axis=1)
    m.species = list(m.species) + ["X"]
    ld = two_level(nu=second_nu, A=second_A, mass=second_mass, num=3, partner_num=2,
                   Cd_cgs=3e-11, T_table=(10.0, 100.0, 1000.0, 3000.0))
    m.more_species = [LineSpecies(ld, m.roots.copy(), m.qweights.copy())]
    m.meta["species"] = "rotor(4) + two-level"
    return m
