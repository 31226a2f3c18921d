"""GPU parity of the banded Lambda* (n_off_diag >= 2) against the oracle.

The band elements (RayPair::update_Lambda, raypair_lambda.tpp:129-146) are
computed by k_lambda_band in FP64 whatever the precision of the radiation
field, so they are held to the FP64 bar (1e-10 relative) against the
oracle's lists, merged per emitting cell as Lambda::add_element keeps them
(lambda.tpp:130-143).  J_eff subtracts every element (sim_lines.tpp:176-220);
the rate equations with the cross-cell elements
(lineProducingSpecies.tpp:317-340) are solved by block Jacobi on the GPU and
by a dense LU in the oracle; one solve from identical inputs agrees to 1e-10,
and whole ALI runs converge to populations within 1e-5 of the oracle's.
"""
import numpy as np
import pytest

from magritte_amd.setup import synthetic
from oracle.py_oracle import Oracle

pytestmark = pytest.mark.gpu

MODELS = {
    # moving envelope, interpolated steps, 3 lines
    "envelope5": lambda: synthetic.envelope_model(n=5, nside=1, nquads=11, nlev=4, seed=3),
    # sub-thermal envelope (the band terms couple most where collisions are weak)
    "subthermal6": lambda: synthetic.envelope_model(n=6, n0=1.0e12, T0=1000.0, nside=1, seed=3,
                                                    abundance=1.0e-3),
    # 48 rays, 7 levels
    "envelope5_48": lambda: synthetic.envelope_model(n=5, nside=2, nquads=9, seed=9),
}


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def _pair(name, n_off_diag):
    from magritte_amd.hip import Context
    m = MODELS[name]()
    m.n_off_diag = n_off_diag
    print(f"\n[band] {name} n_off_diag {n_off_diag}", flush=True)
    return m, Oracle(m).setup(), Context(m).setup()


def _band_err(orc, ctx):
    """Largest relative difference of the merged elements; the sets of
    emitting cells must agree."""
    band_o = orc.lambda_band()
    band_g, diag_add = ctx.lambda_band()
    worst = 0.0
    n = 0
    for p in range(orc.N):
        for k in range(orc.L):
            e = dict(band_o[p][k])
            g = band_g.get((p, k), {})
            assert set(e) == set(g), (p, k, sorted(e), sorted(g))
            for nr, v in e.items():
                worst = max(worst, abs(g[nr] - v) / abs(v))
                n += 1
    return worst, n, diag_add


@pytest.mark.parametrize("name,n_off_diag", [("envelope5", 2), ("envelope5", 4), ("subthermal6", 3),
                                             ("envelope5_48", 2)])
def test_band_elements_match_oracle(name, n_off_diag):
    m, orc, ctx = _pair(name, n_off_diag)
    orc.compute_radiation_field()
    ctx.compute_radiation_field("fp64")
    assert rel_err(ctx.J(), orc.J()) < 1e-10
    worst, n, diag_add = _band_err(orc, ctx)
    print(f"[band] {n} elements, max rel err {worst:.3e}", flush=True)
    assert n > 0
    assert worst < 1e-10
    # the element naming the origin itself is held with the diagonal
    Lo = orc.lambda_diag()
    Lg = ctx.lambda_diag() + diag_add
    mask = Lo != 0
    assert rel_err(Lg[mask], Lo[mask]) < 1e-10
    # J_eff subtracts every element
    orc.compute_Jeff()
    ctx.compute_Jeff()
    _, Je_o, Jd_o = orc.Jeff()
    _, Je_g, Jd_g = ctx.Jeff()
    assert rel_err(Jd_g, Jd_o) < 1e-10
    assert rel_err(Je_g, Je_o) < 1e-10


def test_band_elements_do_not_depend_on_the_field_precision():
    m, orc, ctx = _pair("envelope5", 3)
    ctx.compute_radiation_field("fp32")
    c32, L32 = ctx.lambda_band_raw()
    ctx.compute_radiation_field("fp64")
    c64, L64 = ctx.lambda_band_raw()
    assert np.array_equal(c32, c64)
    assert np.array_equal(L32, L64)


@pytest.mark.parametrize("name,n_off_diag", [("envelope5", 3), ("subthermal6", 2)])
def test_rate_equations_with_bands_match_oracle(name, n_off_diag):
    """One statistical-equilibrium solve from identical J_eff and Lambda*."""
    m, orc, ctx = _pair(name, n_off_diag)
    orc.compute_radiation_field()
    orc.compute_Jeff()
    ctx.compute_radiation_field("fp64")
    ctx.compute_Jeff()
    pop0 = ctx.populations().copy()
    orc.update_using_statistical_equilibrium()
    ctx.compute_level_populations_from_stateq()
    p_o, p_g = orc.populations(), ctx.populations()
    err = rel_err(p_g, p_o)
    print(f"[band] one solve: populations max rel err {err:.3e}", flush=True)
    assert err < 1e-10
    h1, _, _ = ctx.population_history()
    assert np.array_equal(h1, pop0)
    # block Jacobi: a few sweeps, the last one moving nothing above rounding
    n, changes = ctx.band_sweeps()
    print(f"[band] {n} block-Jacobi sweeps, moves {changes}", flush=True)
    assert 2 <= n <= 10
    assert changes[-1] < 1e-14
    tot = m.abundance[:, m.linedata.num]
    assert rel_err(p_g.sum(axis=1), tot) < 1e-12


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_ali_with_bands_matches_oracle(precision):
    m, orc, ctx = _pair("subthermal6", 2)
    it_o = orc.compute_level_populations(True, 60, 1)
    it_g = ctx.compute_level_populations(True, 60, "max", precision)
    err = rel_err(ctx.populations(), orc.populations())
    print(f"[band] ALI {precision}: oracle {it_o} iterations, GPU {it_g}, populations max rel err {err:.3e}",
          flush=True)
    assert err < 1e-5
    assert abs(it_g - it_o) <= 1


def test_bands_are_refused_in_a_batch():
    from magritte_amd.hip import BatchContext, HipError
    ms = []
    for seed in (1, 2):
        m = synthetic.envelope_model(n=4, nside=1, nquads=7, seed=seed)
        m.n_off_diag = 2
        ms.append(m)
    with pytest.raises(HipError, match="batch"):
        BatchContext(ms)
