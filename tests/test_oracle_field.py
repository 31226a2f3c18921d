"""Pins the oracle's whole-model radiation field and convergence statistics
by limits the mathematics fixes.

* Optically thin limit (Simulation::compute_radiation_field,
  sim_radiation.tpp:56-338): with negligible opacity and no velocity field u
  at the origin is the mean of the two incoming boundary intensities, so
  J = sum over ray pairs of 2 w_r (I+ + I-)/2 = B_nu(T_CMB) exactly when the
  weights sum to one.  A lost factor 2 w_r, a ray pair counted once, or a
  boundary intensity read from the wrong side fails it.
* Optically thick LTE limit: in an isothermal static medium in LTE the line
  source function eta/chi is B_nu0(T) at the line frequency (emissivity and
  opacity of lineProducingSpecies.tpp:41-70), and deep inside an optically
  thick medium u = S for bins near line centre; a wrong emissivity/opacity
  prefactor, a missing nu factor in one of them or a wrong LTE ratio fails it.
* check_for_convergence (lineProducingSpecies.tpp:120-171): fraction of
  levels above 1e-10 n_tot whose relative change 2|n - n1|/(n + n1) exceeds
  pop_prec, and the mean change, against numpy on planted populations.
"""
import numpy as np
import pytest

from magritte_amd.setup import synthetic
from magritte_amd.setup.linedata import CC, HH, KB
from oracle.py_oracle import Oracle

T_CMB = 2.7254800


def planck(T, nu):
    return 2.0 * HH * nu ** 3 / CC ** 2 / np.expm1(HH * nu / (KB * T))


def test_optically_thin_limit_gives_cmb():
    m = synthetic.regular_grid_model(n=8, abundance=1.0e-22)
    orc = Oracle(m).setup()
    orc.compute_radiation_field()
    nu, _ = orc.frequencies()
    J = orc.J()
    expect = planck(T_CMB, nu)
    rel = np.abs(J - expect) / expect
    assert rel.max() < 1e-8, rel.max()


def test_optically_thick_lte_gives_line_source_function():
    n = 10
    m = synthetic.regular_grid_model(n=n, abundance=1.0e-1, n_H2=1.0e14, T=40.0)
    orc = Oracle(m).setup()
    orc.compute_radiation_field()
    ld = m.linedata
    nu0 = ld.frequency[0]
    dE = ld.energy[ld.irad[0]] - ld.energy[ld.jrad[0]]
    B0 = 2.0 * HH * nu0 ** 3 / CC ** 2 / np.expm1(dE / (KB * 40.0))
    J = orc.J()
    nu, _ = orc.frequencies()
    # central bins (|x| <= 1 Doppler width) of cells at least 3 cells from the boundary
    i = np.arange(n)
    I, Jg, K = np.meshgrid(i, i, i, indexing="ij")
    deep = ((np.minimum(I, n - 1 - I) >= 3) & (np.minimum(Jg, n - 1 - Jg) >= 3)
            & (np.minimum(K, n - 1 - K) >= 3)).ravel(order="C")
    cells = np.nonzero(deep)[0]
    assert cells.size > 0
    roots = m.roots
    central = np.nonzero(np.abs(roots) <= 1.0)[0]
    vals = J[np.ix_(cells, central)]
    rel = np.abs(vals - B0) / B0
    assert rel.max() < 1e-6, rel.max()
    # and a source function off by a factor shows up: the boundary-adjacent
    # cells see the CMB and fall well below B0 in the wings
    assert J[:, 0].min() < 0.9 * B0


def test_check_for_convergence_statistics():
    m = synthetic.envelope_model(n=4, seed=5)
    orc = Oracle(m).setup()
    rng = np.random.default_rng(0)
    pop = orc.populations()
    N, V = pop.shape
    prev = pop * (1.0 + rng.choice([0.0, 1e-7, 1e-5, 3e-3], size=pop.shape))
    # a few levels below 1e-10 n_tot are ignored whatever their change
    ntot = m.abundance[:, m.linedata.num]
    pop[0, V - 1] = 1e-12 * ntot[0]
    prev[0, V - 1] = 5.0 * pop[0, V - 1]
    orc.set_populations(pop)
    h = orc.population_history()
    orc.set_population_history(prev, h[1], h[2])
    fnc, mean, mx = orc.check_for_convergence(1e-6)
    mask = pop > 1e-10 * ntot[:, None]
    rc = 2.0 * np.abs(pop - prev) / (pop + prev)
    assert fnc == pytest.approx(np.count_nonzero(rc[mask] > 1e-6) / (N * V), rel=1e-12)
    assert mean == pytest.approx(rc[mask].sum() / (N * V), rel=1e-12)
    assert mx == pytest.approx(rc[mask].max(), rel=1e-12)


def test_scattering_u_per_ray_sums_to_J():
    """Radiation::u with scattering (radiation.cpp:117-126, sim_radiation.tpp:226-262):
    the oracle keeps every ray pair's u at the origin; J is the ray-weighted
    sum 2 w_r u_r in ray order, and a pair of one point holds I_bdy."""
    m = synthetic.regular_grid_model(n=6)
    o = Oracle(m).setup()
    o.set_scattering(True)
    o.compute_radiation_field()
    u = o.u()
    w = np.asarray(m.weights)
    J = np.zeros_like(u[0])
    for r in range(m.nrays // 2):
        J = J + 2.0 * w[r] * u[r]
    np.testing.assert_array_equal(J, o.J())
    assert np.all(u > 0)
    o.set_scattering(False)
    o.compute_radiation_field()
    np.testing.assert_array_equal(o.J(), J)


def test_n_off_diag_one_is_the_diagonal_operator():
    """n_off_diag = 1 solves each pair with the back-substitution branch
    (raypair_solver.tpp:336-388) but update_Lambda adds bands from m = 1 on
    (raypair_lambda.tpp:129-146): J and Lambda* equal the n_off_diag = 0 ones to
    rounding.  With n_off_diag = 2, J is the same and the band elements come
    on top (pinned in test_oracle_band.py)."""
    m0 = synthetic.envelope_model(n=6, seed=5)
    o0 = Oracle(m0).setup()
    o0.compute_radiation_field()
    m1 = synthetic.envelope_model(n=6, seed=5)
    m1.n_off_diag = 1
    o1 = Oracle(m1).setup()
    o1.compute_radiation_field()
    np.testing.assert_allclose(o1.J(), o0.J(), rtol=1e-12, atol=0.0)
    np.testing.assert_allclose(o1.lambda_diag(), o0.lambda_diag(), rtol=1e-12, atol=0.0)
    m2 = synthetic.envelope_model(n=6, seed=5)
    m2.n_off_diag = 2
    o2 = Oracle(m2).setup()
    o2.compute_radiation_field()
    np.testing.assert_allclose(o2.J(), o0.J(), rtol=1e-12, atol=0.0)
    assert sum(len(e) for row in o2.lambda_band() for e in row) > 0
    assert sum(len(e) for row in o1.lambda_band() for e in row) == 0
