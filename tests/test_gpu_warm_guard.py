"""Safety of the learned warm start and of the ALI loop against diverging
starts, through the C-ABI.

- mgr_warm_start keeps a cell's previous populations when the network's
  prediction is not finite and positive or inverts a line (opacity <= 0):
  a network whose output inverts every line leaves the LTE start unchanged,
  bit for bit; one whose output inverts none is taken as predicted.  The
  same holds per cell over the concatenated models of a batch.
- The trained network far outside its training data (the optically thick
  stress envelope, n0 = 1e17 m^-3, where its predictions invert lines) still
  converges to the LTE-started populations.
- Non-finite populations stop mgr_compute_level_populations with an error
  instead of passing the criterion (whose comparisons are false for NaN).
"""
import numpy as np
import pytest

from magritte_amd.hip import MLP_HIDDEN, MLP_IN, MLP_OUT
from magritte_amd.setup import synthetic

pytestmark = pytest.mark.gpu


def constant_network(y):
    """A network whose output is y (log10 fractions) for every input."""
    from magritte_amd.hip import Mlp
    net = Mlp(seed=0)
    net.set_params({"W1": np.zeros((MLP_IN, MLP_HIDDEN)), "b1": np.zeros(MLP_HIDDEN),
                    "W2": np.zeros((MLP_HIDDEN, MLP_HIDDEN)), "b2": np.zeros(MLP_HIDDEN),
                    "W3": np.zeros((MLP_HIDDEN, MLP_OUT)), "b3": np.asarray(y, np.float32)})
    net.set_normalisation(np.zeros(MLP_IN), np.ones(MLP_IN), np.zeros(MLP_OUT), np.ones(MLP_OUT))
    return net


def _context(model):
    from magritte_amd.hip import Context
    ctx = Context(model).setup()
    ctx.trace()
    ctx.first_pass("fp32", 1)
    return ctx


def test_inverted_predictions_keep_the_previous_populations():
    m = synthetic.envelope_model(n=6, nside=1, nquads=9, seed=2)
    print("\n[guard] envelope 6^3, constant predictions", flush=True)
    ctx = _context(m)
    P_lte = ctx.populations().copy()
    # 10^(i/2): every upper level holds more per statistical weight than the
    # one below (n_u / n_l = 3.16 > g_u / g_l <= 3 for the rotor): all lines invert
    bad = constant_network(0.5 * np.arange(MLP_OUT))
    assert ctx.warm_start(bad) == m.ncells
    assert np.array_equal(ctx.populations(), P_lte)
    # 10^-i: no inversion, taken as predicted
    y = -np.arange(MLP_OUT, dtype=np.float32)
    good = constant_network(y)
    assert ctx.warm_start(good) == 0
    f = 10.0 ** y.astype(np.float64)
    expect = m.abundance[:, m.linedata.num][:, None] * (f / f.sum())[None, :]
    P = ctx.populations()
    assert float(np.max(np.abs(P - expect) / expect)) < 1e-12
    # the ALI from the accepted start reaches the LTE-started populations
    it = ctx.compute_level_populations(True, 60, "max", "fp32")
    ref = _context(m)
    it_lte = ref.compute_level_populations(True, 60, "max", "fp32")
    err = float(np.max(np.abs(ctx.populations() - ref.populations()) / ref.populations()))
    print(f"[guard] from 10^-i: {it} iterations, LTE {it_lte}; populations differ by {err:.2e}", flush=True)
    assert err < 1e-5
    for c in (ctx, ref):
        c.close()
    bad.close()
    good.close()


def test_trained_network_outside_its_data_still_converges():
    from magritte_amd import warmstart
    m = synthetic.stress_model(n=16)
    print("\n[guard] stress 16^3 (n0 1e17), trained network", flush=True)
    net = warmstart.load(warmstart.DEFAULT_WEIGHTS)
    ctx = _context(m)
    ref = _context(m)
    it_lte = ref.compute_level_populations(True, 60, "max", "fp32")
    if net.first_pass_stride > 0:
        ctx.first_pass("fp32", net.first_pass_stride)
    n_fb = ctx.warm_start(net)
    P0 = ctx.populations()
    assert np.all(np.isfinite(P0)) and np.all(P0 > 0)
    it = ctx.compute_level_populations(True, 60, "max", "fp32")
    P, P_ref = ctx.populations(), ref.populations()
    assert np.all(np.isfinite(P))
    err = float(np.max(np.abs(P - P_ref) / P_ref))
    print(f"[guard] {n_fb} of {m.ncells} cells kept LTE; {it} iterations (LTE {it_lte}); "
          f"populations differ by {err:.2e}", flush=True)
    assert err < 1e-5
    for c in (ctx, ref):
        c.close()
    net.close()


def test_nonfinite_populations_stop_the_iteration():
    from magritte_amd.hip import Context, HipError
    m = synthetic.envelope_model(n=5, nside=1, nquads=9, seed=4)
    print("\n[guard] envelope 5^3 with a NaN population", flush=True)
    ctx = Context(m).setup()
    P = ctx.populations().copy()
    P[17, 3] = np.nan
    ctx.set_populations(P)
    with pytest.raises(HipError, match="non-finite"):
        ctx.compute_level_populations(True, 10, "max", "fp32")
    ctx.close()


def test_guard_in_a_batch():
    """The guard per cell over the concatenated models of a batch."""
    from magritte_amd.hip import BatchContext
    ms = [synthetic.envelope_model(n=n, nside=1, nquads=9, seed=s) for n, s in ((5, 1), (6, 2))]
    print("\n[guard] batch of 5^3 and 6^3 envelopes", flush=True)
    b = BatchContext(ms).setup()
    b.trace()
    b.first_pass("fp32", 1)
    P_lte = b.populations().copy()
    bad = constant_network(0.5 * np.arange(MLP_OUT))
    assert b.warm_start(bad) == b.N
    assert np.array_equal(b.populations(), P_lte)
    good = constant_network(-np.arange(MLP_OUT, dtype=np.float32))
    assert b.warm_start(good) == 0
    assert not np.array_equal(b.populations(), P_lte)
    b.close()
    bad.close()
    good.close()
