"""The tracer's paths against the oracle's traces, bit-exact
(Geometry::trace_ray, geometry.tpp:21-95, and get_next_general,
geometry.tpp:328-421, as the oracle restates them):

- candidate masks on and off (the masks skip the neighbours behind the
  current cell along the ray);
- the counting pass's record replayed by the filling pass for every walk,
  for none (every walk scanned again), and for a record of 4 cells (nearly
  every walk longer than it, so filled by the second scan);
- neighbour lists longer than the masks hold (63 entries): scanned whole;
- walks longer than the default record (256 cells), on an elongated grid.

Every setting must also give bit-identical radiation fields and the same
window-skip statistics (which read the whole cache, padding included).
"""
import numpy as np
import pytest

from magritte_amd.model import Model
from magritte_amd.setup import synthetic
from magritte_amd.setup.linedata import CC, two_level
from magritte_amd.setup.quadrature import gaussian_linspace
from oracle.py_oracle import Oracle

pytestmark = pytest.mark.gpu

OPTIONS = [(-1, True), (0, True), (4, True), (-1, False), (4, False)]


def _grid(nx, ny, nz):
    p = np.arange(nx * ny * nz)
    return p % nx, (p // nx) % ny, p // (nx * ny)


def _neighbours(nx, ny, nz, radius):
    """Lists of every cell within Chebyshev distance `radius`, dz outer, dy,
    dx inner."""
    i, j, k = _grid(nx, ny, nz)
    offs = [(dx, dy, dz) for dz in range(-radius, radius + 1) for dy in range(-radius, radius + 1)
            for dx in range(-radius, radius + 1) if (dx, dy, dz) != (0, 0, 0)]
    cols = []
    for dx, dy, dz in offs:
        ii, jj, kk = i + dx, j + dy, k + dz
        ok = (ii >= 0) & (ii < nx) & (jj >= 0) & (jj < ny) & (kk >= 0) & (kk < nz)
        cols.append(np.where(ok, ii + nx * jj + nx * ny * kk, -1))
    nb = np.stack(cols, axis=1)
    return (nb >= 0).sum(axis=1).astype(np.int64), nb[nb >= 0].astype(np.int64)


def long_ray_model(nx=300, seed=4):
    """nx x 3 x 3 grid, 26-neighbour lists, the six axis rays: the walks along
    x from the centre line run up to nx - 2 steps.  A sinusoidal x velocity
    makes some steps interpolated."""
    rng = np.random.default_rng(seed)
    ny = nz = 3
    h = 1.0e12
    i, j, k = _grid(nx, ny, nz)
    N = nx * ny * nz
    pos = np.stack([(i + 0.5) * h, (j + 0.5) * h, (k + 0.5) * h], axis=1).astype(np.float64)
    pos += rng.uniform(-0.05, 0.05, size=pos.shape) * h
    vel = np.zeros((N, 3))
    vel[:, 0] = 2.0e-5 * np.sin(2.0 * np.pi * i / 40.0)
    n_nb, nbs = _neighbours(nx, ny, nz, 1)
    bdy = np.nonzero((i == 0) | (i == nx - 1) | (j == 0) | (j == ny - 1) | (k == 0) | (k == nz - 1))[0]
    rays = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1]], dtype=np.float64)
    roots, qw = gaussian_linspace(11, 5.0)
    ab = np.zeros((N, 3))
    ab[:, 1] = 1.0e-4 * 1.0e12
    ab[:, 2] = 1.0e12
    return Model(position=pos, velocity=vel, n_neighbors=n_nb, neighbors=nbs, rays=rays,
                 weights=np.full(6, 1.0 / 6.0), boundary2cell_nr=bdy.astype(np.int64),
                 temperature=np.full(N, 20.0), vturb2=np.full(N, (150.0 / CC) ** 2), abundance=ab,
                 species=["dummy", "CO", "H2"], linedata=two_level(), roots=roots, qweights=qw)


def wide_list_model():
    """6^3 envelope whose neighbour lists hold every cell within two grid
    steps (up to 124 entries)."""
    m = synthetic.envelope_model(n=6, seed=11, nside=1, nquads=9)
    m.n_neighbors, m.neighbors = _neighbours(6, 6, 6, 2)
    return m


def _compare_rays(model, orc, ctx, origins):
    checked = 0
    for o in origins:
        for r in range(model.nrays):
            _, c_o, s_o, z_o = orc.trace_ray(int(o), r)
            c_g, s_g, z_g = ctx.ray(int(o), r)
            assert np.array_equal(c_o, c_g), (o, r)
            assert np.array_equal(s_o, s_g), (o, r)
            assert np.array_equal(z_o, z_g), (o, r)
            checked += len(c_o)
    return checked


def _fingerprint(ctx):
    ctx.compute_radiation_field("fp32")
    return ctx.J().copy(), ctx.window_skip_stats(), ctx.npoints_on_rays().copy()


def test_every_option_bit_exact_envelope():
    from magritte_amd.hip import Context
    m = synthetic.envelope_model(n=8, seed=7)
    orc = Oracle(m).setup()
    ctx = Context(m).setup()
    origins = np.random.default_rng(5).choice(m.ncells, size=24, replace=False)
    ref = None
    for rec, masks in OPTIONS:
        print(f"\n[trace] envelope8 record {rec} masks {masks}", flush=True)
        ctx.set_trace_options(rec, masks)
        ctx.trace()
        assert _compare_rays(m, orc, ctx, origins) > 0
        fp = _fingerprint(ctx)
        if ref is None:
            ref = fp
        else:
            assert np.array_equal(fp[0], ref[0])
            assert fp[1] == ref[1]
            assert np.array_equal(fp[2], ref[2])
    ctx.close()


def test_lists_longer_than_the_masks():
    from magritte_amd.hip import Context
    m = wide_list_model()
    assert int(m.n_neighbors.max()) > 63
    print("\n[trace] wide lists", int(m.n_neighbors.max()), flush=True)
    orc = Oracle(m).setup()
    ctx = Context(m).setup()
    ctx.trace()
    assert _compare_rays(m, orc, ctx, range(0, m.ncells, 7)) > 0
    J0 = _fingerprint(ctx)[0]
    ctx.set_trace_options(0, True)
    ctx.trace()
    assert np.array_equal(_fingerprint(ctx)[0], J0)
    ctx.close()


def test_walks_longer_than_the_record():
    from magritte_amd.hip import Context
    m = long_ray_model()
    print("\n[trace] elongated grid", m.ncells, "cells", flush=True)
    orc = Oracle(m).setup()
    ctx = Context(m).setup()
    ctx.trace()
    centre = [i + 300 * (1 + 3 * 1) for i in (1, 2, 40, 150, 297, 298)]   # the j = k = 1 line
    assert max(orc.npoints_on_ray(o, 0) for o in centre) > 256
    assert _compare_rays(m, orc, ctx, centre) > 0
    fp0 = _fingerprint(ctx)
    ctx.set_trace_options(1024, True)   # every walk recorded
    ctx.trace()
    assert _compare_rays(m, orc, ctx, centre[:2]) > 0
    fp1 = _fingerprint(ctx)
    assert np.array_equal(fp0[0], fp1[0]) and fp0[1] == fp1[1]
    ctx.close()


def test_option_arguments_checked():
    from magritte_amd.hip import Context, HipError
    ctx = Context(synthetic.envelope_model(n=4, nside=1, nquads=7)).setup()
    for bad in (-2, 6, 8192):
        with pytest.raises(HipError, match="record_steps"):
            ctx.set_trace_options(bad, True)
    ctx.close()
