"""GPU parity of the ALI loop through Ng acceleration, against the FP64 oracle.

Simulation::compute_level_populations (src/Simulation/sim_lines.tpp:61-164)
runs Ng acceleration (LineProducingSpecies::update_using_Ng_acceleration,
lineProducingSpecies.tpp:181-238) after every four normal iterations, so a
model must need at least 10 iterations for two Ng steps to run.  The model
here is an optically thin-to-moderate, sub-thermal envelope (10^3 cells,
n0 = 1e12 m^-3, T0 = 1000 K, abundance 1e-3, 12 rays) that the oracle takes 11
iterations to converge at max relative change < 1e-6.

  * one Ng step on identical populations and history (taken from four normal
    oracle iterations): populations, history shift, convergence statistics and
    emissivity/opacity within 1e-12;
  * the whole ALI loop: the same iteration count as the oracle, populations
    within 1e-5 (FP32 radiation field) and 1e-9 (FP64);
  * the FP64 trajectory iteration by iteration, Ng steps included: populations
    within 1e-9 and the mean relative change within 1e-6 of the oracle's after
    every iteration.
"""
import numpy as np
import pytest

from magritte_amd.setup import synthetic
from oracle.py_oracle import Oracle

pytestmark = pytest.mark.gpu

MAX_CRITERION = 1   # MGR_CRIT_MAX


def slow_model():
    return synthetic.envelope_model(n=10, n0=1.0e12, T0=1000.0, nside=1, seed=3, abundance=1.0e-3)


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def _ctx(model):
    from magritte_amd.hip import Context
    return Context(model).setup()


def _oracle_iteration(orc, ng):
    """One iteration of sim_lines.tpp:101-132 on the oracle."""
    if ng:
        orc.update_using_Ng_acceleration()
    else:
        orc.compute_radiation_field()
        orc.compute_Jeff()
        orc.update_using_statistical_equilibrium()
    conv = orc.check_for_convergence()
    orc.set_emissivity_and_opacity()
    return conv


@pytest.fixture(scope="module")
def oracle_run():
    m = slow_model()
    print("\n[ali] oracle ALI on the 10^3 sub-thermal envelope", flush=True)
    orc = Oracle(m).setup()
    it = orc.compute_level_populations(True, 60, MAX_CRITERION)
    return m, it, orc.populations()


def test_model_needs_two_ng_steps(oracle_run):
    _, it, _ = oracle_run
    assert it >= 10, it


def test_one_ng_step_matches_oracle():
    m = slow_model()
    orc = Oracle(m).setup()
    for _ in range(4):
        _oracle_iteration(orc, False)
    pop = orc.populations()
    h1, h2, h3 = orc.population_history()
    print("[ali] one Ng step from four oracle iterations", flush=True)
    ctx = _ctx(m)
    ctx.set_populations(pop)
    ctx.set_population_history(h1, h2, h3)
    conv_g = ctx.iteration_using_Ng_acceleration()
    conv_o = _oracle_iteration(orc, True)
    P_o, P_g = orc.populations(), ctx.populations()
    assert not np.array_equal(P_o, pop), "the Ng step must change the populations"
    assert rel_err(P_g, P_o) < 1e-12
    for a, b in zip(ctx.population_history(), orc.population_history()):
        assert np.array_equal(a, b)
    assert np.array_equal(ctx.population_history()[0], pop)
    # the reference sums the weight 1/(ncells nlev) per unconverged level under an
    # OpenMP reduction, so the fraction's last bits depend on the summation order
    assert abs(conv_g[0] - conv_o[0]) <= 1e-12 * max(conv_o[0], 1.0 / P_o.size)
    assert abs(conv_g[1] - conv_o[1]) <= 1e-12 * conv_o[1]
    assert abs(conv_g[2] - conv_o[2]) <= 1e-12 * conv_o[2]
    e_o, k_o = orc.emissivity_opacity()
    e_g, k_g = ctx.emissivity_opacity()
    assert rel_err(e_g, e_o) < 1e-12
    assert np.max(np.abs(k_g - k_o) / np.maximum(np.abs(k_o), 1e-300)) < 1e-11
    ctx.close()


def test_ng_step_with_zero_determinant_changes_nothing():
    """A history with no change (all four iterates equal) gives D = 0: the
    reference skips the update and the history shift (lineProducingSpecies.tpp:220)."""
    m = synthetic.envelope_model(n=6, seed=5)
    orc = Oracle(m).setup()
    pop = orc.populations()
    ctx = _ctx(m)
    ctx.set_populations(pop)
    ctx.set_population_history(pop, pop, pop)
    ctx.iteration_using_Ng_acceleration()
    assert np.array_equal(ctx.populations(), pop)
    for a in ctx.population_history():
        assert np.array_equal(a, pop)
    ctx.close()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_ali_through_ng_matches_oracle(oracle_run, precision, tol):
    m, it_o, P_o = oracle_run
    ctx = _ctx(m)
    it_g = ctx.compute_level_populations(True, 60, "max", precision)
    err = rel_err(ctx.populations(), P_o)
    print(f"[ali] {precision}: oracle {it_o} iterations, GPU {it_g}, populations max rel err {err:.3e}",
          flush=True)
    assert it_g == it_o
    assert err < tol
    ctx.close()


def test_ali_trajectory_fp64_matches_oracle(oracle_run):
    m, it_o, _ = oracle_run
    orc = Oracle(m).setup()
    ctx = _ctx(m)
    iteration_normal = 0
    ng_steps = 0
    for it in range(1, it_o + 1):
        ng = iteration_normal == 4
        conv_o = _oracle_iteration(orc, ng)
        conv_g = ctx.ali_iteration(ng, "fp64")
        iteration_normal = 0 if ng else iteration_normal + 1
        ng_steps += int(ng)
        err = rel_err(ctx.populations(), orc.populations())
        print(f"[ali] iteration {it} ({'Ng' if ng else 'normal'}): pops max rel err {err:.3e}, "
              f"mean change {conv_g[1]:.6e} vs {conv_o[1]:.6e}", flush=True)
        assert err < 1e-9, it
        assert abs(conv_g[1] - conv_o[1]) <= 1e-6 * conv_o[1], it
        assert abs(conv_g[2] - conv_o[2]) <= 1e-6 * conv_o[2] + 1e-14, it
    assert ng_steps >= 2
    ctx.close()
