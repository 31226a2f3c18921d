"""Pins the oracle's line physics, set-up and level-population update.

* get_eta_and_chi (simulation.tpp:385-481): numpy brute force over all lines
  with the +-11-width window, the 1e-26 floor (P3) and the rule that the last
  line is reached when the frequency lies above every window (P2).
* LTE populations (lineProducingSpecies.tpp:82-115), Planck boundary
  intensities (planck.tpp:15-24, sim_radiation.tpp:17-48) and the spectral
  grid (sim_spectrum.tpp:13-196): closed forms.
* Collision rates (collisionPartner.tpp:15-72): linear interpolation, constant
  below the first table temperature, linear extrapolation above the last, and
  the ortho/para H2 fraction 1/(1 + 9 exp(-170.5/T)).
* Statistical equilibrium (lineProducingSpecies.tpp:249-495): two-level closed
  form n_u/n_l = (B_a J + C_e)/(A + B_s J - L* kappa + C_d), and for the
  7-level rotor the balance of in- and out-flow of every level computed from
  the physical rates (the detailed-balance invariant of
  tests/unit/Model/Lines/test_lineProducingSpecies.cpp:71-173).
* Ng acceleration (lineProducingSpecies.tpp:181-238): exact for sequences
  with two geometric modes.
* J_eff (sim_lines.tpp:176-220): closed-form quadrature sum.
"""
import numpy as np
import pytest

from magritte_amd.setup import synthetic
from magritte_amd.setup.linedata import CC, HH, KB
from magritte_amd.setup.quadrature import gauss_hermite
from oracle.py_oracle import Oracle, planck

AMU = 1.66053904E-27
C_SQUARED = 8.98755179E+16
HH_OVER_FOUR_PI = HH / (4.0 * np.pi)


def profile_width(m, p):
    return np.sqrt(2.0 * KB / (AMU * C_SQUARED) * m.linedata.inverse_mass * m.temperature[p] + m.vturb2[p])


def brute_eta_chi(m, emis, opac, freq, p):
    """All lines l with -11 w <= freq - nu_l <= 11 w, plus the last line when
    freq lies above every window."""
    lines = np.sort(m.linedata.frequency)
    order = np.argsort(m.linedata.frequency, kind="stable")
    pw = profile_width(m, p)
    eta, chi = 0.0, 1.0e-26
    skipped_all = True
    for l, nu_l in enumerate(lines):
        w = nu_l * pw
        fd = freq - nu_l
        if fd <= 11.0 * w:
            skipped_all = False
        if -11.0 * w <= fd <= 11.0 * w:
            phi = 1.0 / (w * np.sqrt(np.pi)) * np.exp(-(fd / w) ** 2)
            eta += emis[p, order[l]] * freq * phi
            chi += opac[p, order[l]] * freq * phi
    if skipped_all:
        l = len(lines) - 1
        w = lines[l] * pw
        fd = freq - lines[l]
        phi = 1.0 / (w * np.sqrt(np.pi)) * np.exp(-(fd / w) ** 2)
        eta += emis[p, order[l]] * freq * phi
        chi += opac[p, order[l]] * freq * phi
    return eta, chi


@pytest.fixture(scope="module", params=["envelope", "blended"])
def env(request):
    # blended: groups of 2 and 4 lines about one width apart (several lines in a window)
    m = (synthetic.envelope_model(n=6, seed=2) if request.param == "envelope"
         else synthetic.blended_model(n=5, seed=2))
    orc = Oracle(m).setup()
    return m, orc


def test_eta_chi_window(env):
    m, orc = env
    emis, opac = orc.emissivity_opacity()
    rng = np.random.default_rng(0)
    lines = np.sort(m.linedata.frequency)
    for p in rng.choice(m.ncells, 10, replace=False):
        p = int(p)
        w = lines * profile_width(m, p)
        freqs = np.concatenate([lines + w * rng.uniform(-12, 12, (4, len(lines))).ravel().reshape(4, -1),
                                ]).ravel()
        freqs = np.concatenate([freqs, [lines[0] * 0.5, lines[-1] * 1.5]])
        for fr in freqs:
            e_o, c_o, _ = orc.eta_and_chi(float(fr), p)
            e_b, c_b = brute_eta_chi(m, emis, opac, float(fr), p)
            assert e_o == pytest.approx(e_b, rel=1e-12, abs=1e-300)
            assert c_o == pytest.approx(c_b, rel=1e-12)


def test_last_line_rule_and_floor():
    """Above every window the last line is still included; below every window
    only the 1e-26 floor remains.  Huge opacities make the exp(-x^2 > 121)
    term visible in FP64."""
    m = synthetic.regular_grid_model(n=4, nquads=5)
    m.abundance[:, 1] = 1e200           # enormous line-species density
    orc = Oracle(m).setup()
    emis, opac = orc.emissivity_opacity()
    nu = m.linedata.frequency[0]
    w = nu * profile_width(m, 0)
    e_hi, c_hi, _ = orc.eta_and_chi(nu + 12.0 * w, 0)
    assert c_hi > 1.0e-26
    e_b, c_b = brute_eta_chi(m, emis, opac, nu + 12.0 * w, 0)
    assert c_hi == pytest.approx(c_b, rel=1e-12)
    e_lo, c_lo, _ = orc.eta_and_chi(nu - 12.0 * w, 0)
    assert c_lo == 1.0e-26 and e_lo == 0.0


def test_lte_and_emissivity(env):
    m, orc = env
    ld = m.linedata
    pops = orc.populations()
    for p in range(0, m.ncells, 17):
        w = ld.weight * np.exp(-ld.energy / (KB * m.temperature[p]))
        expect = w * (m.abundance[p, ld.num] / w.sum())
        assert np.allclose(pops[p], expect, rtol=1e-14)
    emis, opac = orc.emissivity_opacity()
    i, j = ld.irad, ld.jrad
    assert np.allclose(emis, HH_OVER_FOUR_PI * ld.A * pops[:, i], rtol=1e-14)
    assert np.allclose(opac, HH_OVER_FOUR_PI * (pops[:, j] * ld.Ba - pops[:, i] * ld.Bs), rtol=1e-12)


def test_spectral_grid_and_boundary_intensities(env):
    m, orc = env
    nu, nr = orc.frequencies()
    lines = np.sort(m.linedata.frequency)
    for p in range(0, m.ncells, 13):
        pw = profile_width(m, p)
        raw = np.array([lines[k] + (lines[k] * pw) * m.roots[z] for k in range(m.nrad) for z in range(m.nquads)])
        assert np.array_equal(np.sort(raw), nu[p])
        for k in range(m.nrad):
            for z in range(m.nquads):
                assert nu[p, nr[p, k, z]] == raw[k * m.nquads + z]
    I = orc.boundary_intensities()
    T = 2.72548
    for b in range(0, m.nboundary, 7):
        f = nu[m.boundary2cell_nr[b]]
        expect = 2.0 * HH / C_SQUARED * f ** 3 / np.expm1(HH / KB / T * f)
        assert np.allclose(I[b], expect, rtol=1e-13)
    assert planck(20.0, 1e11) == pytest.approx(2 * HH / C_SQUARED * 1e33 / np.expm1(HH / KB / 20.0 * 1e11), rel=1e-14)


def test_collision_rates():
    m = synthetic.envelope_model(n=4)
    orc = Oracle(m)
    cp = m.linedata.colpar[0]
    for T in (5.0, cp.tmp[0], 33.3, 500.0, cp.tmp[-1], 5000.0):
        ce, cd, fac = orc.collision_rates(0, T)
        t = np.searchsorted(cp.tmp, T)
        if T <= cp.tmp[0]:
            exp_cd = cp.Cd[0]
        else:
            t = min(max(t, 1), len(cp.tmp) - 1)
            step = (T - cp.tmp[t - 1]) / (cp.tmp[t] - cp.tmp[t - 1])
            exp_cd = cp.Cd[t - 1] + (cp.Cd[t] - cp.Cd[t - 1]) * step
        assert np.allclose(cd, exp_cd, rtol=1e-12, atol=0)
        assert fac == 1.0
    cp.orth_or_para = "p"
    orc2 = Oracle(m)
    _, _, fac = orc2.collision_rates(0, 50.0)
    assert fac == pytest.approx(1.0 / (1.0 + 9.0 * np.exp(-170.5 / 50.0)))
    cp.orth_or_para = "o"
    orc3 = Oracle(m)
    _, _, fac = orc3.collision_rates(0, 50.0)
    assert fac == pytest.approx(1.0 - 1.0 / (1.0 + 9.0 * np.exp(-170.5 / 50.0)))


def test_two_level_statistical_equilibrium():
    m = synthetic.regular_grid_model(n=6, nquads=11)
    m.abundance[:, 2] = 1.0e8          # sub-thermal: radiation matters
    orc = Oracle(m).setup()
    pops0 = orc.populations()
    orc.compute_radiation_field()
    orc.compute_Jeff()
    _, Jeff, _ = orc.Jeff()
    Ls = orc.lambda_diag()
    orc.update_using_statistical_equilibrium()
    pops = orc.populations()
    ld = m.linedata
    cp = ld.colpar[0]
    Cd = cp.Cd[0][0] * m.abundance[:, 2]
    Ce = cp.Ce[0][0] * m.abundance[:, 2]
    kappa = HH_OVER_FOUR_PI * (pops0[:, 0] * ld.Ba[0] - pops0[:, 1] * ld.Bs[0])
    ratio = (ld.Ba[0] * Jeff[:, 0] + Ce) / (ld.A[0] + ld.Bs[0] * Jeff[:, 0] - Ls[:, 0] * kappa + Cd)
    assert np.allclose(pops[:, 1] / pops[:, 0], ratio, rtol=1e-10)
    assert np.allclose(pops.sum(axis=1), m.abundance[:, 1], rtol=1e-12)
    # a sign error in the ALO term would change the ratio measurably
    wrong = (ld.Ba[0] * Jeff[:, 0] + Ce) / (ld.A[0] + ld.Bs[0] * Jeff[:, 0] + Ls[:, 0] * kappa + Cd)
    assert not np.allclose(pops[:, 1] / pops[:, 0], wrong, rtol=1e-6)


def test_multilevel_balance(env):
    """After the stat-eq update every level's in-flow equals its out-flow with
    the physical rates (radiative A, B J_eff minus the ALO term, collisional)."""
    m, _ = env
    orc = Oracle(m).setup()
    pops0 = orc.populations()
    orc.compute_radiation_field()
    orc.compute_Jeff()
    _, Jeff, _ = orc.Jeff()
    Ls = orc.lambda_diag()
    orc.update_using_statistical_equilibrium()
    n = orc.populations()
    ld = m.linedata
    V = ld.nlev
    cp = ld.colpar[0]
    for p in range(0, m.ncells, 11):
        R = np.zeros((V, V))     # R[i, j]: rate i -> j per particle in i
        for k in range(ld.nrad):
            i, j = ld.irad[k], ld.jrad[k]
            kap = HH_OVER_FOUR_PI * (pops0[p, j] * ld.Ba[k] - pops0[p, i] * ld.Bs[k])
            R[i, j] += ld.A[k] + ld.Bs[k] * Jeff[p, k] - Ls[p, k] * kap
            R[j, i] += ld.Ba[k] * Jeff[p, k]
        ce, cd, _ = orc.collision_rates(0, m.temperature[p])
        abn = m.abundance[p, cp.num]
        for k in range(len(cp.icol)):
            i, j = cp.icol[k], cp.jcol[k]
            R[i, j] += cd[k] * abn
            R[j, i] += ce[k] * abn
        inflow = n[p] @ R
        outflow = n[p] * R.sum(axis=1)
        scale = np.max(np.abs(n[p][:, None] * R))
        assert np.allclose(inflow[:V - 1], outflow[:V - 1], rtol=0, atol=1e-9 * scale)
        assert n[p].sum() == pytest.approx(m.abundance[p, ld.num], rel=1e-12)


def test_ng_exact_for_two_modes(env):
    m, _ = env
    orc = Oracle(m).setup()
    rng = np.random.default_rng(5)
    N, V = m.ncells, m.nlev
    xs = rng.uniform(1, 2, (N, V))
    a1, a2 = rng.uniform(-0.1, 0.1, (2, N, V))
    l1, l2 = 0.7, -0.4
    seq = [xs + a1 * l1 ** k + a2 * l2 ** k for k in range(4)]
    # prev3 = x_0, prev2 = x_1, prev1 = x_2, population = x_3
    orc.set_population_history(seq[2], seq[1], seq[0])
    orc.set_populations(seq[3])
    orc.update_using_Ng_acceleration()
    assert np.allclose(orc.populations(), xs, rtol=1e-9)
    p1, p2, p3 = orc.population_history()
    assert np.allclose(p1, seq[3]) and np.allclose(p2, seq[2]) and np.allclose(p3, seq[1])


def test_jeff_quadrature_sum(env):
    m, _ = env
    orc = Oracle(m).setup()
    orc.compute_radiation_field()
    orc.compute_Jeff()
    J = orc.J()
    _, nr = orc.frequencies()
    Jlin, Jeff, Jdif = orc.Jeff()
    Ls = orc.lambda_diag()
    pops = orc.populations()
    expect = (m.qweights[None, None, :] * np.take_along_axis(J[:, None, :].repeat(m.nrad, 1), nr, 2)).sum(axis=2)
    assert np.allclose(Jlin, expect, rtol=1e-14)
    assert np.allclose(Jdif, HH_OVER_FOUR_PI * Ls * pops[:, m.linedata.irad], rtol=1e-14)
    assert np.allclose(Jeff, Jlin - Jdif, rtol=1e-12, atol=0)


def test_gauss_hermite_normalised():
    r, w = gauss_hermite(7)
    assert w.sum() == pytest.approx(1.0, rel=1e-13)
    rr, ww = np.polynomial.hermite.hermgauss(7)
    assert np.allclose(r, rr) and np.allclose(w, ww / np.sqrt(np.pi))
    assert (w * r ** 2).sum() == pytest.approx(0.5, rel=1e-12)   # <x^2> of exp(-x^2)/sqrt(pi)
