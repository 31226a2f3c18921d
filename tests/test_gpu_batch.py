"""A batch of independent models in one context (mgr_create_batch) against
the same models solved one by one.

Bars: traced rays bit-exact (the tracer never leaves a model); J within 1e-6
relative (FP32 kernel) and 1e-12 (FP64 kernel) of the single-model solves;
per-model iteration counts equal and converged populations within 1e-9 of
the single-model runs (Ng and the convergence test are per model, and a
converged model is frozen).
"""
import numpy as np
import pytest

from magritte_amd.setup import synthetic

pytestmark = pytest.mark.gpu


def _models():
    return [synthetic.envelope_model(n=7, n0=1e13, T0=800.0, nside=1, seed=31),
            synthetic.envelope_model(n=9, n0=1e15, T0=2500.0, nside=2, seed=32),
            synthetic.envelope_model(n=6, n0=1e12, T0=1500.0, nside=2, seed=33),
            synthetic.envelope_model(n=8, n0=1e14, T0=3000.0, nside=1, seed=34)]


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


@pytest.fixture(scope="module")
def batch():
    from magritte_amd.hip import BatchContext, Context
    print("\n[batch] 4 envelopes (7^3/12, 9^3/48, 6^3/48, 8^3/12 rays)", flush=True)
    ms = _models()
    b = BatchContext(ms).setup()
    b.trace()
    singles = []
    for m in ms:
        c = Context(m).setup()
        c.trace()
        singles.append(c)
    return ms, b, singles


def test_batch_layout(batch):
    ms, b, _ = batch
    assert b.M == 4
    assert list(np.diff(b.cell_offsets)) == [m.ncells for m in ms]


def test_batch_rays_bit_exact(batch):
    ms, b, singles = batch
    rng = np.random.default_rng(0)
    for k, (m, c) in enumerate(zip(ms, singles)):
        off = b.cell_offsets[k]
        for o in rng.choice(m.ncells, size=5, replace=False):
            for r in (0, m.nrays // 2 - 1, m.nrays - 1):
                cb, sb, zb = b.ray(int(off + o), r)
                cs, ss, zs = c.ray(int(o), r)
                assert np.array_equal(cb - off, cs)
                assert np.array_equal(sb, ss) and np.array_equal(zb, zs)


@pytest.mark.parametrize("prec,tol", [("fp32", 1e-6), ("fp64", 1e-12)])
def test_batch_radiation_field(batch, prec, tol):
    ms, b, singles = batch
    b.compute_radiation_field(prec)
    Jb = b.split(b.J())
    Lb = b.split(b.lambda_diag())
    for k, c in enumerate(singles):
        c.compute_radiation_field(prec)
        assert _rel(Jb[k], c.J()) < tol
        L = c.lambda_diag()
        mask = L > 0
        assert _rel(Lb[k][mask], L[mask]) < tol


def test_batch_level_populations(batch):
    ms, b, singles = batch
    from magritte_amd.hip import BatchContext, Context
    b2 = BatchContext(ms).setup()
    b2.trace()
    b2.compute_level_populations(True, 200, "max", "fp32")
    its, conv = b2.convergence()
    P = b2.split(b2.populations())
    for k, m in enumerate(ms):
        c = Context(m).setup()
        c.trace()
        it = c.compute_level_populations(True, 200, "max", "fp32")
        print(f"[batch] model {k}: iterations batch {its[k]} single {it}", flush=True)
        assert its[k] == it
        assert _rel(P[k], c.populations()) < 1e-9
        assert conv[k][2] < m.pop_prec
        c.close()
    b2.close()
