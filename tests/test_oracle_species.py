"""Pins the oracle's handling of several line-producing species
(Parameters::nlspecs > 1).

* Decoupling: a rotor with lines at 115-346 GHz and a two-level species at
  700 GHz, of equal mass.  Their lines are more than 11 widths apart, so each
  bin's eta and chi see one species only, both species give the same
  dshift_max and the merged line list keeps each species' lines in order.  The
  two-species model must then reproduce, bin for bin and level for level, the
  two single-species models: J, Lambda*, J_eff and six ALI iterations with an
  Ng step (sim_lines.tpp:61-164; Lines::iteration_using_* loops the species).
* get_dshift_max (simulation.tpp:15-37) is the smallest width over species.
* The reference's own quirks, restated:
  - compute_spectral_discretisation (sim_spectrum.tpp:23-45) centres species
    l's line k on the merged sorted list at the running line index and uses
    species l's width; with interleaved lines the bins of a species sit on
    another species' line;
  - get_line_width (simulation.tpp:363-373) takes the species from the
    frequency back-map at the index line_index[lindex], which is a bin index:
    every line whose merged index is below nrad_0 * nquads gets species 0's
    mass.
  A numpy restatement of each is checked against the oracle, and a planted
  change (the species' own line, the species' own mass) fails.
"""
import copy

import numpy as np
import pytest

from magritte_amd.setup import synthetic
from magritte_amd.setup.linedata import CC, KB
from oracle.py_oracle import Oracle

AMU = 1.66053904E-27
C_SQUARED = 8.98755179E+16


def _single(m, l):
    """Species l of m as a one-species model."""
    s = copy.deepcopy(m)
    if l > 0:
        sp = m.more_species[l - 1]
        s.linedata, s.roots, s.qweights = sp.linedata, sp.roots, sp.qweights
    s.more_species = []
    return s


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def test_two_species_decouple_into_single_species_runs():
    """Static and isothermal, so that every cell has the same grid: equal up
    to rounding."""
    m = synthetic.two_species_model(n=5, seed=5, pop_prec=1e-30)
    m.velocity[:] = 0.0
    m.temperature[:] = 300.0
    tol = 1e-13
    o2 = Oracle(m).setup()
    oa, ob = Oracle(_single(m, 0)).setup(), Oracle(_single(m, 1)).setup()
    FA, LA = oa.F, oa.L
    for o in (o2, oa, ob):
        o.compute_radiation_field()
        o.compute_Jeff()
    J2 = o2.J()
    assert rel(J2[:, :FA], oa.J()) < tol and rel(J2[:, FA:], ob.J()) < tol
    L2 = o2.lambda_diag()
    assert rel(L2[:, :LA], oa.lambda_diag()) < tol and rel(L2[:, LA:], ob.lambda_diag()) < tol
    _, Je2, _ = o2.Jeff()
    assert rel(Je2[:, :LA], oa.Jeff()[1]) < tol and rel(Je2[:, LA:], ob.Jeff()[1]) < tol
    for o in (o2, oa, ob):
        assert o.compute_level_populations(True, 6, 0) == 6
    pa, pb = o2.populations()
    assert rel(pa, oa.populations()) < 10 * tol and rel(pb, ob.populations()) < 10 * tol


def test_boundary_intensity_is_interpolated_across_species():
    """In the moving envelope a far end's shifted frequency below the far
    cell's first bin of the upper species lands between the two species' bins
    of the merged grid, and Radiation::rescale_I_bdy (radiation.tpp:341-373)
    interpolates between them, where the single-species run takes its first
    bin.  Away from that edge (the upper half of the species' bins) the two
    agree."""
    m = synthetic.two_species_model(n=5, seed=5)
    o2 = Oracle(m).setup()
    oa, ob = Oracle(_single(m, 0)).setup(), Oracle(_single(m, 1)).setup()
    for o in (o2, oa, ob):
        o.compute_radiation_field()
    FA, Q = oa.F, m.nquads
    J2 = o2.J()
    assert rel(J2[:, :FA], oa.J()) < 1e-9
    upper = slice(FA + Q // 2 + 1, FA + Q)
    assert rel(J2[:, upper], ob.J()[:, Q // 2 + 1:]) < 1e-9
    assert rel(J2[:, FA:FA + 3], ob.J()[:, :3]) > 1e-6
    # the driver's convergence figures are the largest over the species
    cs = o2.convergence_species()
    fnc, rcm, rcmax = o2.convergence()
    assert fnc == max(c[0] for c in cs) and rcmax == max(c[2] for c in cs)
    assert cs[0][2] == pytest.approx(oa.convergence()[2], rel=1e-10)


def _pw(T, vturb2, mass):
    return np.sqrt(2.0 * KB / (AMU * C_SQUARED) * (1.0 / mass) * T + vturb2)


def test_dshift_max_is_the_smallest_over_species():
    m = synthetic.two_species_model(n=4, second_mass=2.0)
    o = Oracle(m)
    for p in (0, 17, 40):
        want = m.max_width_fraction * min(_pw(m.temperature[p], m.vturb2[p], 28.0),
                                          _pw(m.temperature[p], m.vturb2[p], 2.0))
        assert o.dshift_max(p) == pytest.approx(want, rel=1e-14)


def spectral_grid(m, lines_sorted):
    """sim_spectrum.tpp:23-45: bins of species l, line k centred on the sorted
    list at the running line index, with species l's width and roots."""
    species = [(m.linedata, m.roots)] + [(s.linedata, s.roots) for s in m.more_species]
    out = np.zeros((m.ncells, m.nfreqs))
    for p in range(m.ncells):
        f, idx0 = [], 0
        for ld, roots in species:
            pw = _pw(m.temperature[p], m.vturb2[p], 1.0 / ld.inverse_mass)
            for k in range(ld.nrad):
                line = lines_sorted[idx0]
                f.extend(line + line * pw * roots)
                idx0 += 1
        out[p] = np.sort(np.array(f))
    return out


@pytest.mark.parametrize("second_nu", [700.0e9, 200.0e9])
def test_spectral_grid_of_several_species(second_nu):
    m = synthetic.two_species_model(n=4, second_nu=second_nu, second_mass=5.0)
    o = Oracle(m)
    o.compute_spectral_discretisation()
    nu, _ = o.frequencies()
    lines, _ = o.lines()
    assert np.all(np.diff(lines) > 0)
    assert rel(nu, spectral_grid(m, lines)) < 1e-14
    # back-maps by unsorted position: species 0's bins first
    appears, l, k, z = o.line_maps()
    assert np.all(appears == 1)
    FA = m.linedata.nrad * m.nquads
    assert np.all(l[:FA] == 0) and np.all(l[FA:] == 1)
    if second_nu < m.linedata.frequency.max():
        # planted: each species on its own line frequency gives another grid
        own = m.species_linedata
        lines_own = np.concatenate([ld.frequency for ld in own])
        assert rel(nu, spectral_grid(m, lines_own)) > 1e-6


def test_line_width_takes_the_species_of_the_bin_back_map():
    """With species 1 lighter, its line's width in eta/chi is species 0's
    (the back-map at bin index line_index[lindex] < nrad_0 * nquads)."""
    m = synthetic.two_species_model(n=4, second_mass=2.0)
    o = Oracle(m).setup()
    ld1 = m.more_species[0].linedata
    nu1 = ld1.frequency[0]
    emis, opac = o.emissivity_opacity()
    g = m.linedata.nrad                      # species 1's line in the merged list
    for p in (0, 21):
        eta, chi, _ = o.eta_and_chi(nu1, p, 0)
        width0 = nu1 * _pw(m.temperature[p], m.vturb2[p], 28.0)
        prof = 1.0 / width0 / np.sqrt(np.pi)
        assert chi == pytest.approx(1e-26 + opac[p, g] * nu1 * prof, rel=1e-12)
        assert eta == pytest.approx(emis[p, g] * nu1 * prof, rel=1e-12)
        width1 = nu1 * _pw(m.temperature[p], m.vturb2[p], 2.0)   # planted: the species' own mass
        assert abs(chi - (1e-26 + opac[p, g] * nu1 / width1 / np.sqrt(np.pi))) > 1e-3 * chi
