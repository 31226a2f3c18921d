"""The out-of-window skip and the far-end intensity cache of the
radiation-field kernels against the full evaluation, through the C-ABI.

A step whose sub-points keep every bin of a 64-bin wave more than 10.5 line
widths from every line has FP32 profiles of exactly 0 (exp2 of less than
-149), so eta = 0 and chi sits at its floor; the skip eliminates such a step
without evaluating the profiles.  The FP64 kernel has bits of its own, set
beyond 11 (1 + 1e-6) widths, where get_eta_and_chi's +-11-width window
(simulation.tpp:385-481) holds no line: eta = 0 and chi at the floor there
too.  Bar: J, the Lambda diagonal, Radiation::u and the populations after ALI
iterations are bit-identical with the skip on and off, in both precisions,
and a model with velocity fields does skip a share of its points.
The masks follow the frequencies: after an image discretisation and back the
line-grid result is unchanged.  The far-end intensity cache (the rescaled
boundary intensity of every ray and bin, kept in FP32) gives the same bits as
the kernel's own search, and is rebuilt when the boundary intensities change.
"""
import numpy as np
import pytest

from magritte_amd.setup import synthetic

pytestmark = pytest.mark.gpu


def _jump_model():
    m = synthetic.regular_grid_model(n=6)
    v = np.zeros((m.ncells, 3))
    v[m.position[:, 0] > 0.5 * m.position[:, 0].max(), 0] = 8.4e-4
    m.velocity = v
    return m


MODELS = {
    "envelope12": (lambda: synthetic.envelope_model(n=12, seed=5), True),
    "stress8": (lambda: synthetic.stress_model(n=8, seed=11), True),
    "regular8": (lambda: synthetic.regular_grid_model(n=8), False),
    "jump6": (_jump_model, False),
    "blended8": (lambda: synthetic.blended_model(n=8, seed=21), False),
}


def _field(ctx, precision="fp32"):
    ctx.compute_radiation_field(precision)
    return ctx.J(), ctx.lambda_diag(), ctx.u()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("name", list(MODELS))
def test_skip_is_bit_identical(name, precision):
    from magritte_amd.hip import Context
    make, must_skip = MODELS[name]
    print(f"\n[skip] model {name} {precision}", flush=True)
    ctx = Context(make()).setup()
    ctx.set_scattering(True)
    J1, L1, u1 = _field(ctx, precision)
    total, skipped = ctx.window_skip_stats()
    print(f"[skip] {name}: {skipped} of {total} point-chunks skipped ({skipped / max(total, 1):.1%})", flush=True)
    assert 0 <= skipped <= total
    if must_skip:
        assert skipped > 0
    ctx.set_window_skip(False)
    J0, L0, u0 = _field(ctx, precision)
    assert np.array_equal(J1, J0)
    assert np.array_equal(L1, L0)
    assert np.array_equal(u1, u0)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_skip_in_ali_iterations(precision):
    from magritte_amd.hip import Context
    print(f"\n[skip] envelope10 ALI {precision}", flush=True)
    ctx = Context(synthetic.envelope_model(n=10, seed=9)).setup()
    pops = {}
    for on in (True, False):
        ctx.set_window_skip(on)
        ctx.set_far_cache(on)
        ctx.compute_LTE_level_populations()
        for _ in range(3):
            ctx.ali_iteration(ng=False, precision=precision)
        pops[on] = ctx.populations()
    assert np.array_equal(pops[True], pops[False])


def test_far_cache_is_bit_identical_and_follows_the_boundary():
    from magritte_amd.hip import Context
    print("\n[skip] stress8 far-end intensity cache", flush=True)
    m = synthetic.stress_model(n=8, seed=11)
    ctx = Context(m).setup()
    ctx.set_scattering(True)
    J1, L1, u1 = _field(ctx)
    ctx.set_far_cache(False)
    J0, L0, u0 = _field(ctx)
    assert np.array_equal(J1, J0) and np.array_equal(L1, L0) and np.array_equal(u1, u0)
    # new boundary temperatures: the cache is rebuilt from them
    ctx.set_far_cache(True)
    ctx.compute_boundary_intensities(np.full(m.nboundary, 40.0))
    J2, _, _ = _field(ctx)
    ctx.set_far_cache(False)
    J3, _, _ = _field(ctx)
    assert np.array_equal(J2, J3)
    assert not np.array_equal(J2, J1)


def test_masks_follow_the_frequencies():
    from magritte_amd.hip import Context
    print("\n[skip] envelope10 line -> image -> line grid", flush=True)
    ctx = Context(synthetic.envelope_model(n=10, seed=9)).setup()
    ctx.compute_radiation_field("fp32")
    J1 = ctx.J()
    ctx.compute_spectral_discretisation_image(3.0e-5)
    ctx.compute_boundary_intensities()
    Im, Ip = ctx.compute_image(0, "fp32")
    assert np.all(np.isfinite(Im)) and np.all(np.isfinite(Ip))
    ctx.compute_spectral_discretisation()
    ctx.compute_boundary_intensities()
    ctx.compute_radiation_field("fp32")
    assert np.array_equal(ctx.J(), J1)


def test_far_cache_fp64_is_bit_identical():
    from magritte_amd.hip import Context
    print("\n[skip] stress8 far-end intensity cache, FP64 kernel", flush=True)
    ctx = Context(synthetic.stress_model(n=8, seed=11)).setup()
    ctx.compute_radiation_field("fp64")
    J1, L1 = ctx.J(), ctx.lambda_diag()
    ctx.set_far_cache(False)
    ctx.compute_radiation_field("fp64")
    assert np.array_equal(ctx.J(), J1) and np.array_equal(ctx.lambda_diag(), L1)
