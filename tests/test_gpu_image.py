"""GPU images against the FP64 oracle, through the C-ABI
(Simulation::compute_and_write_image, sim_radiation.tpp:345-478;
RayPair::solve_for_image, raypair_solver.tpp:523-579;
compute_spectral_discretisation_image, sim_spectrum.tpp:206-311;
get_npoints_on_rays<Rest>, simulation.tpp:540-552).

Bars:
  * image frequency grid and nr_line: bit-exact;
  * rest-frame point counts per ray pair: exact;
  * I_m, I_p: 1e-10 relative for the FP64 kernel, on LTE populations and on
    those after one ALI iteration.  For the FP32 kernel (FP32 sweep and
    profiles, FP64 shifts and profile arguments) 2e-5 of the intensities the
    result is formed from, |I| + 2 I_bdy (one ray's intensity: no average over
    ray pairs as in J, whose bar is 1e-5; measured up to 1.03e-5 on the
    stress model, 4e-6 on the envelope).  I = 2u - I_bdy cancels where the
    medium absorbs the boundary intensity (the reference's own result is
    negative in ~0.3 % of the stress model's bins), so an error relative to I
    alone measures that cancellation, not the solve;
  * the state checks of the reference: an image needs the image grid, the
    level populations need the line grid.
"""
import numpy as np
import pytest

from magritte_amd.setup import synthetic
from oracle.py_oracle import Oracle

pytestmark = pytest.mark.gpu

I_TOL = {"fp64": 1e-10, "fp32": 2e-5}
WIDTH = 3.0e-5   # image bins span +-(width x max root) of each line frequency

MODELS = {
    "regular16": lambda: synthetic.regular_grid_model(n=16),
    "envelope8": lambda: synthetic.envelope_model(n=8, seed=7),
    "stress8": lambda: synthetic.stress_model(n=8, seed=11),
}


def rel_err(a, b, scale=None):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.abs(b) if scale is None else np.abs(b) + scale
    return float(np.max(np.abs(a - b) / np.maximum(den, 1e-300)))


@pytest.fixture(scope="module", params=list(MODELS))
def image_pair(request):
    from magritte_amd.hip import Context
    print(f"\n[image] model {request.param}", flush=True)
    model = MODELS[request.param]()
    # populations: LTE and after one oracle ALI iteration on the line grid
    orc = Oracle(model).setup()
    lte = orc.populations()
    orc.compute_radiation_field()
    orc.compute_Jeff()
    orc.update_using_statistical_equilibrium()
    nlte = orc.populations()
    orc.compute_spectral_discretisation_image(WIDTH)
    orc.compute_boundary_intensities()
    ctx = Context(model)
    ctx.compute_spectral_discretisation_image(WIDTH)
    ctx.compute_boundary_intensities()
    ctx.compute_LTE_level_populations()
    return request.param, model, orc, ctx, {"lte": lte, "nlte": nlte}


def test_image_grid_bit_exact(image_pair):
    _, model, orc, ctx, _ = image_pair
    nu_o, nr_o = orc.frequencies()
    nu_g, nr_g = ctx.frequencies()
    assert np.array_equal(nu_o, nu_g)
    assert np.array_equal(nr_o, nr_g)
    assert rel_err(ctx.boundary_intensities(), orc.boundary_intensities()) < 1e-14


def test_npoints_on_rays_rest(image_pair):
    _, model, orc, ctx, _ = image_pair
    npg = ctx.npoints_on_rays_rest()
    ant = orc.antipod()
    for r in range(model.nrays // 2):
        for o in range(0, model.ncells, max(1, model.ncells // 61)):
            n_r = len(orc.trace_ray(o, r, rest=True)[1])
            n_a = len(orc.trace_ray(o, int(ant[r]), rest=True)[1])
            assert npg[r, o] == n_r + n_a + 1, (r, o)


@pytest.mark.parametrize("pops", ["lte", "nlte"])
def test_image_matches_oracle(image_pair, pops):
    name, model, orc, ctx, pp = image_pair
    orc.set_populations(pp[pops])
    ctx.set_populations(pp[pops])
    # the boundary intensity of each bin (every boundary cell holds the same CMB grid)
    Ib2 = 2.0 * orc.boundary_intensities().max(axis=0)[None, :]
    for r in sorted({0, model.nrays // 4, model.nrays // 2 - 1}):
        Im_o, Ip_o = orc.compute_image(r)
        for prec in ("fp64", "fp32"):
            print(f"[image] {name} {pops} ray {r} {prec}", flush=True)
            Im_g, Ip_g = ctx.compute_image(r, prec)
            assert np.all(np.isfinite(Im_g)) and np.all(np.isfinite(Ip_g))
            sc = None if prec == "fp64" else Ib2
            em, ep = rel_err(Im_g, Im_o, sc), rel_err(Ip_g, Ip_o, sc)
            print(f"[image]   I_m {em:.2e}  I_p {ep:.2e}  (plain relative {rel_err(Im_g, Im_o):.1e}, "
                  f"{rel_err(Ip_g, Ip_o):.1e})", flush=True)
            assert em < I_TOL[prec] and ep < I_TOL[prec], (r, prec, em, ep)


def test_image_needs_image_grid_and_lines_need_line_grid():
    from magritte_amd.hip import Context, HipError
    model = synthetic.regular_grid_model(n=6)
    ctx = Context(model).setup()
    with pytest.raises(HipError, match="Images"):
        ctx.compute_image(0)
    ctx.compute_spectral_discretisation_image(WIDTH)
    ctx.compute_boundary_intensities()
    with pytest.raises(HipError, match="Lines"):
        ctx.compute_level_populations(max_niterations=2)
    with pytest.raises(HipError, match="nrays/2"):
        ctx.compute_image(model.nrays // 2)
    Im, Ip = ctx.compute_image(1)
    assert np.all(Im > 0) and np.all(Ip > 0)
    ctx.compute_spectral_discretisation()
    ctx.compute_boundary_intensities()
    ctx.compute_level_populations(max_niterations=2)
