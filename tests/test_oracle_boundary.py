"""Pins the oracle's boundary-intensity rescaling and its Doppler-shift step
limit against independent numpy restatements.

* Radiation::rescale_I_bdy (src/Model/Radiation/radiation.tpp:341-373) with
  search_with_notch from notch 0 (src/Functions/interpolation.tpp:96-110) and
  interpolate_linear (:126-134).  The rule, restated independently here:
  notch = the first bin whose frequency is >= freq_scaled (np.searchsorted,
  side "left"), clipped to the last bin; notch 0 or the last bin returns that
  bin's intensity unchanged; otherwise linear interpolation between bins
  notch-1 and notch.  Checked at every node, at interior midpoints (against
  np.interp), below the first bin, above the last bin, and between the last two
  bins, where the reference returns the last bin's value and np.interp would
  interpolate: a wrong neighbour, a wrong edge rule or a swapped end fails.
* The same rescaling seen through the whole radiation field: on a moving,
  optically thin model, u at the origin is the mean of the two incoming
  intensities, so J(o, f) = sum over ray pairs of w_r (I_ar + I_r), with each I
  the rescaled boundary intensity of the side's last cell at that side's last
  shift times nu(o, f) in that cell's own frequency grid, or the origin's own
  boundary intensity for an empty side (simulation.tpp:225-353).  A shift taken
  from the wrong point, the origin's grid instead of the far cell's, or the
  wrong boundary index fails it.
* Simulation::get_dshift_max (src/Simulation/simulation.tpp:15-37) with
  Thermodynamics::profile_width (thermodynamics.tpp:62-69):
  max_width_fraction * sqrt(2 k_B T / (amu c^2) / m + vturb^2), c^2 the
  reference's literal C_SQUARED (src/Tools/constants.hpp:45).  It is then fed
  into the tracer's interpolation count, int(|dshift| / dshift_max + 1)
  (geometry.tpp:100-207), with the independent value.
"""
import numpy as np

from magritte_amd.setup import synthetic
from magritte_amd.setup.linedata import AMU, KB
from oracle.py_oracle import Oracle

C_SQUARED = 8.98755179E+16   # src/Tools/constants.hpp:45


def rescale_numpy(nu, I, v):
    """Independent restatement of rescale_I_bdy's rule for one grid."""
    F = nu.shape[0]
    notch = min(int(np.searchsorted(nu, v, side="left")), F - 1)
    if notch == 0 or notch == F - 1:
        return I[notch]
    x1, x2, f1, f2 = nu[notch], nu[notch - 1], I[notch], I[notch - 1]
    return (f2 - f1) / (x2 - x1) * (v - x1) + f1


def test_rescale_I_bdy_rule():
    m = synthetic.envelope_model(n=6, seed=9)
    orc = Oracle(m).setup()
    nu, _ = orc.frequencies()
    Ib = orc.boundary_intensities()
    F = m.nfreqs
    for p in m.boundary2cell_nr[::17]:
        p = int(p)
        b = orc.cell2boundary_nr(p)
        g, I = nu[p], Ib[b]
        assert np.all(np.diff(g) > 0)
        # every node returns its own bin
        for k in range(F):
            assert orc.rescale_I_bdy(g[k], 0, p) == I[k]
        # interior midpoints: linear interpolation between the two bins
        for k in range(1, F - 1):
            v = 0.5 * (g[k - 1] + g[k])
            got = orc.rescale_I_bdy(v, 0, p)
            assert abs(got - np.interp(v, g, I)) <= 1e-12 * abs(I).max()
            assert got == rescale_numpy(g, I, v)
        # a quarter of the way from one bin to the next: not symmetric, so a
        # swapped pair of bins fails
        for k in range(1, F - 1):
            v = g[k - 1] + 0.25 * (g[k] - g[k - 1])
            want = I[k - 1] + 0.25 * (I[k] - I[k - 1])
            assert abs(orc.rescale_I_bdy(v, 0, p) - want) <= 1e-12 * abs(I).max()
        # edges: below the grid, above it, and between the last two bins
        assert orc.rescale_I_bdy(g[0] * (1 - 1e-4), 0, p) == I[0]
        assert orc.rescale_I_bdy(g[-1] * (1 + 1e-4), 0, p) == I[-1]
        v = 0.5 * (g[-2] + g[-1])
        assert orc.rescale_I_bdy(v, 0, p) == I[-1]
        assert I[-1] != np.interp(v, g, I)


def test_thin_moving_model_boundary_terms():
    m = synthetic.envelope_model(n=6, seed=9, abundance=1.0e-22)
    orc = Oracle(m).setup()
    orc.compute_radiation_field()
    J = orc.J()
    nu, _ = orc.frequencies()
    Ib = orc.boundary_intensities()
    ant = orc.antipod()
    w = np.asarray(m.weights)
    b_of = {int(p): k for k, p in enumerate(m.boundary2cell_nr)}
    rng = np.random.default_rng(1)
    interpolated = 0
    for o in rng.choice(m.ncells, 40, replace=False):
        o = int(o)
        expect = np.zeros(m.nfreqs)
        for r in range(m.nrays // 2):
            for side in (r, int(ant[r])):
                _, cells, shifts, _ = orc.trace_ray(o, side)
                if len(cells) == 0:
                    expect += w[r] * Ib[b_of[o]]
                    continue
                c = int(cells[-1])
                for f in range(m.nfreqs):
                    v = shifts[-1] * nu[o, f]
                    expect[f] += w[r] * rescale_numpy(nu[c], Ib[b_of[c]], v)
                    interpolated += int(nu[c, 0] < v < nu[c, -2])
        rel = np.abs(J[o] - expect) / expect
        assert rel.max() < 1e-8, (o, rel.max())
    # the rays do reach the interpolating branch, not only the edge bins
    assert interpolated > 1000


def test_dshift_max_closed_form():
    for m in (synthetic.envelope_model(n=6, seed=9), synthetic.regular_grid_model(n=6)):
        orc = Oracle(m)
        inv_mass = m.linedata.inverse_mass
        width = np.sqrt(2.0 * KB / (AMU * C_SQUARED) * inv_mass * m.temperature + m.vturb2)
        expect = m.max_width_fraction * width
        got = np.array([orc.dshift_max(o) for o in range(m.ncells)])
        np.testing.assert_allclose(got, expect, rtol=1e-15, atol=0.0)
    # the parameter is honoured, not a constant 0.5
    m = synthetic.envelope_model(n=4, seed=9)
    m.max_width_fraction = 0.2
    orc = Oracle(m)
    width = np.sqrt(2.0 * KB / (AMU * C_SQUARED) * m.linedata.inverse_mass * m.temperature + m.vturb2)
    np.testing.assert_allclose([orc.dshift_max(o) for o in range(m.ncells)], 0.2 * width, rtol=1e-15)


def test_interpolation_count_uses_dshift_max():
    """The tracer's set_data count int(|dshift| / dshift_max + 1) with the
    closed-form dshift_max: the oracle's point counts on a model with velocity
    jumps follow it (geometry.tpp:100-207)."""
    m = synthetic.stress_model(n=8, seed=4)
    orc = Oracle(m)
    inv_mass = m.linedata.inverse_mass
    dsm = m.max_width_fraction * np.sqrt(2.0 * KB / (AMU * C_SQUARED) * inv_mass * m.temperature + m.vturb2)
    rng = np.random.default_rng(2)
    counted = 0
    for o in rng.choice(m.ncells, 30, replace=False):
        o = int(o)
        for r in range(0, m.nrays, 5):
            crt, cells, shifts, _ = orc.trace_ray(o, r, dshift_max=float(dsm[o]))
            n = m.rays[r]
            prev_shift = 1.0
            total = 0
            i = 0
            while i < len(cells):
                j = i
                if crt[i] < 0:
                    while j < len(cells) and crt[j] == crt[i]:
                        j += 1
                    j -= 1
                c = int(cells[j])
                dv = m.velocity[c] - m.velocity[o]
                s_c = 1.0 - ((dv[0] * n[0] + dv[1] * n[1]) + dv[2] * n[2])
                npts = int(abs(s_c - prev_shift) / dsm[o] + 1)
                assert (j - i + 1) == (npts if npts >= 2 else 1)
                total += npts if npts >= 2 else 1
                counted += int(npts >= 2)
                prev_shift = s_c
                i = j + 1
            assert total == len(cells)
    assert counted > 0
