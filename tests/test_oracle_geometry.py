"""Pins the oracle's geometry (tracer, Doppler shifts, rays) and set-up.

* Doppler known answers of tests/unit/Model/Geometry/test_geometry.cpp:18-66.
* Tracer on a regular grid: axis-aligned rays follow the grid line, dZ is the
  cell size, the ray ends on the first boundary cell (brute force).
* Tracer invariants on the jittered envelope model: Z strictly increasing,
  consecutive cells are neighbours, the ray ends on a boundary cell, point
  counts equal get_npoints_on_ray, plain points carry the Doppler shift
  1 - (v_c - v_o).n of their cell, interpolation points obey set_data's count
  n = floor(|ds|/ds_max + 1).
* HEALPix RING directions: known pixel centres, unit norm, antipodes exist,
  and every reduced ray's antipode lies in the other half.
"""
import math

import numpy as np
import pytest

from magritte_amd.model import Model
from magritte_amd.setup import healpix, synthetic
from magritte_amd.setup.linedata import CC, two_level
from magritte_amd.setup.quadrature import gaussian_linspace
from oracle.py_oracle import Oracle


def line_model(vx, rays):
    n = len(vx)
    pos = np.stack([np.arange(n, dtype=float), np.zeros(n), np.zeros(n)], axis=1)
    vel = np.stack([np.asarray(vx, float), np.zeros(n), np.zeros(n)], axis=1)
    nbs, nn = [], []
    for i in range(n):
        lst = [j for j in (i - 1, i + 1) if 0 <= j < n]
        nbs += lst
        nn.append(len(lst))
    roots, qw = gaussian_linspace(3)
    ab = np.zeros((n, 3)); ab[:, 1] = 1e6; ab[:, 2] = 1e10
    return Model(position=pos, velocity=vel, n_neighbors=np.array(nn), neighbors=np.array(nbs),
                 rays=np.asarray(rays, float), weights=np.full(len(rays), 1.0 / len(rays)),
                 boundary2cell_nr=np.array([0, n - 1]), temperature=np.full(n, 20.0),
                 vturb2=np.full(n, (150 / CC) ** 2), abundance=ab, species=["d", "CO", "H2"],
                 linedata=two_level(), roots=roots, qweights=qw)


def test_doppler_known_answers():
    m = line_model([0.2, 0.3, 0.1, -0.1], [[1.0, 0, 0], [-1.0, 0, 0]])
    orc = Oracle(m)
    como = [orc.doppler_shift(0, 0, c) for c in range(4)]
    rest = [orc.doppler_shift(0, 0, c, rest=True) for c in range(4)]
    assert como == pytest.approx([1.0, 0.9, 1.1, 1.3], abs=1e-15)
    assert rest == pytest.approx([0.8, 0.7, 0.9, 1.1], abs=1e-15)
    assert como[0] == 1.0


def test_axis_aligned_rays_follow_grid_lines():
    n = 7
    m = synthetic.regular_grid_model(n=n, side=7.0e14, nside=1, nquads=3)
    # replace the HEALPix rays by the six axis directions (antipodal pairs)
    m.rays = np.array([[1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0], [-1.0, 0, 0], [0, -1.0, 0], [0, 0, -1.0]])
    m.weights = np.full(6, 1.0 / 6)
    orc = Oracle(m)
    h = 1.0e14
    for (i, j, k) in [(3, 3, 3), (1, 2, 5), (5, 5, 1)]:
        o = i + n * j + n * n * k
        for r, (di, dj, dk) in enumerate([(1, 0, 0), (0, 1, 0), (0, 0, 1), (-1, 0, 0), (0, -1, 0), (0, 0, -1)]):
            _, cells, shift, dZ = orc.trace_ray(o, r)
            expect = []
            ii, jj, kk = i, j, k
            while True:
                ii, jj, kk = ii + di, jj + dj, kk + dk
                if not (0 <= ii < n and 0 <= jj < n and 0 <= kk < n):
                    break
                expect.append(ii + n * jj + n * n * kk)
                if ii in (0, n - 1) or jj in (0, n - 1) or kk in (0, n - 1):
                    break
            assert list(cells) == expect, (o, r)
            assert np.allclose(dZ, h, rtol=1e-14)
            assert np.all(shift == 1.0)


def Rv_step(m, o, a, b, n):
    """Projected distance between cells a and b along n, seen from origin o."""
    za = np.dot(m.position[a] - m.position[o], n) if a != o else 0.0
    zb = np.dot(m.position[b] - m.position[o], n)
    return zb - za


def _neighbour_sets(m):
    off = np.concatenate([[0], np.cumsum(m.n_neighbors)])
    return [set(m.neighbors[off[p]:off[p + 1]].tolist()) for p in range(m.ncells)]


@pytest.mark.parametrize("model", ["envelope", "stress"])
def test_tracer_invariants(model):
    m = synthetic.envelope_model(n=10, seed=3) if model == "envelope" else synthetic.stress_model(n=10, seed=4)
    orc = Oracle(m)
    nbs = _neighbour_sets(m)
    bset = set(m.boundary2cell_nr.tolist())
    rng = np.random.default_rng(0)
    interpolated = 0
    for o in rng.choice(m.ncells, 40, replace=False):
        o = int(o)
        dsm = orc.dshift_max(o)
        for r in range(m.nrays):
            crt, cells, shift, dZ = orc.trace_ray(o, r)
            assert len(cells) == orc.npoints_on_ray(o, r)
            if len(cells) == 0:
                assert o in bset
                continue
            assert cells[-1] in bset
            n = m.rays[r]
            prev_cell, prev_shift, Z = o, 1.0, 0.0
            i = 0
            while i < len(cells):
                if crt[i] >= 0:                    # plain point
                    c = int(cells[i])
                    assert c in nbs[prev_cell]
                    dv = m.velocity[c] - m.velocity[o]
                    assert shift[i] == 1.0 - ((dv[0] * n[0] + dv[1] * n[1]) + dv[2] * n[2])
                    assert abs(shift[i] - prev_shift) <= dsm
                    Rv = m.position[c] - m.position[o]
                    Znew = (Rv[0] * n[0] + Rv[1] * n[1]) + Rv[2] * n[2]
                    assert Znew > Z and dZ[i] > 0
                    Z = Znew
                    prev_cell, prev_shift = c, shift[i]
                    i += 1
                else:                              # the interpolation points of one step
                    j = i
                    while j < len(cells) and crt[j] == crt[i]:
                        j += 1
                    nxt = int(cells[j - 1])
                    assert nxt in nbs[prev_cell]
                    dv = m.velocity[nxt] - m.velocity[o]
                    s_nxt = 1.0 - ((dv[0] * n[0] + dv[1] * n[1]) + dv[2] * n[2])
                    npts = int(abs(s_nxt - prev_shift) / dsm + 1)
                    assert (j - i) == npts >= 2
                    half = int(0.5 * npts)
                    assert np.all(cells[i:i + half - 1] == prev_cell)
                    assert np.all(cells[i + max(half - 1, 0):j] == nxt)
                    assert np.allclose(dZ[i:j].sum(), Rv_step(m, o, prev_cell, nxt, n), rtol=1e-12)
                    assert shift[j - 1] == pytest.approx(s_nxt, abs=1e-15)
                    assert np.all(np.diff(shift[i:j]) * np.sign(shift[j - 1] - shift[i]) >= -1e-15)
                    interpolated += 1
                    Rv = m.position[nxt] - m.position[o]
                    Z = (Rv[0] * n[0] + Rv[1] * n[1]) + Rv[2] * n[2]
                    prev_cell, prev_shift = nxt, shift[j - 1]
                    i = j
    if model == "stress":
        assert interpolated > 0


def test_healpix_ring_pixels():
    for nside in (1, 2, 4):
        rays, w = healpix.ring_rays(nside)
        assert rays.shape == (12 * nside * nside, 3)
        assert np.allclose(np.linalg.norm(rays, axis=1), 1.0, atol=1e-15)
        assert np.allclose(rays.sum(axis=0), 0.0, atol=1e-12)
        assert w.sum() == pytest.approx(1.0)
        ant = healpix.antipodes(rays)
        assert np.all(ant >= 0)
        half = len(rays) // 2
        assert set(ant[:half].tolist()) == set(range(half, len(rays)))
    x, y, z = healpix.pix2vec_ring(1, 0)
    assert z == pytest.approx(2.0 / 3.0)
    assert math.atan2(y, x) == pytest.approx(math.pi / 4)
    x, y, z = healpix.pix2vec_ring(1, 4)
    assert (x, y, z) == pytest.approx((1.0, 0.0, 0.0), abs=1e-15)
    x, y, z = healpix.pix2vec_ring(2, 0)
    assert z == pytest.approx(1.0 - 1.0 / 12.0)


def test_oracle_antipodes_match_setup():
    m = synthetic.envelope_model(n=4)
    assert np.array_equal(Oracle(m).antipod(), healpix.antipodes(m.rays))
