// mgr_radiation.hip -- radiation field: line emissivity/opacity along the
// traced ray pairs and the second-order Feautrier solve, fused in one kernel.
//
// Replaces Simulation::compute_radiation_field (src/Simulation/sim_radiation.tpp:56-338)
// with its inner calls Simulation::setup / get_eta_and_chi (simulation.tpp:185-481),
// Radiation::rescale_I_bdy (radiation.tpp:341-373), RayPair::solve
// (raypair_solver.tpp:200-334, n_off_diag = 0) and RayPair::update_Lambda
// (raypair_lambda.tpp:94-149).
//
// Work mapping (CDNA4, 64-wide wavefronts): one wavefront per (origin cell,
// chunk of 64 frequency bins); lane = frequency bin.  The wavefront loops over
// the nrays/2 ray pairs of its origin; the traced steps and per-cell line data
// are uniform across lanes.  J and the Lambda* diagonal accumulate in FP64
// registers over rays and are written once per (cell, bin): no atomics.
//
// The tridiagonal Feautrier system of a ray pair is solved two-sided: each ray
// of the pair is eliminated from its boundary inwards in the Rybicki-Hummer
// (1+F) form and both eliminations meet at the origin, which yields u and the
// diagonal of the inverse there without storing or back-substituting the pair.
// With a = Dtau_{n-1}, b = Dtau_n the interior row reads, per sweep step,
//   r = b/a, c = (a+b) b / 2, F_n = c + r F_{n-1}/(1+F_{n-1}),
//   Z_n = (S_n c + r Z_{n-1}) / (1+F_n)
// (the reference's F and Su recurrences divided through by C_n), the boundary
// row F_0 = D(1 + D/2), Z_0 = (S D^2/2 + I D)/(1+F_0), and at the origin with
// both sides present, h = (D_A + D_B)/2,
//   u = (S h + Z_A/D_A + Z_B/D_B) / (h + f_A/D_A + f_B/D_B),  f = F/(1+F),
//   Lambda = h / (h + f_A/D_A + f_B/D_B);
// with one side empty (origin on the boundary, incoming I_bdy of the origin's
// boundary cell) u = (S D^2/2 + I D + Z_A)/(D^2/2 + D + f_A),
// Lambda = (D^2/2)/(D^2/2 + D + f_A).  Only ratios and products of optical
// depths appear, so the FP32 variant neither overflows nor cancels.
#include <cfloat>

#include "mgr_common.hpp"
#include "mgr_ray.hpp"

using namespace mgr;

namespace {

constexpr int32_t ERR_NOT_BOUNDARY = 2;   // ray ends on a non-boundary cell
constexpr int32_t ERR_EMPTY_SIDE   = 4;   // empty ray at a non-boundary origin

template <typename real> struct Prec;
template <> struct Prec<float> {
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ float expneg(float x2) { return __expf(-x2); }
};
template <> struct Prec<double> {
    static __device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
    static __device__ __forceinline__ double expneg(double x2) { return exp(-x2); }
};

template <typename real>
struct RFArgs {
    int64_t N;
    int     R, F, L, nchunk;
    const LineTables *tb;
    const double  *nu;                        // [N*F]
    const double  *W11;                       // [N*L] 11 * width
    const real    *IW, *E, *K;                // [N*L] 1/width, scaled prefactors
    const int64_t *seg_off;
    const int32_t *nseg, *npts, *seg_cell;
    const double  *seg_shift, *seg_dZ;
    const int32_t *cell2bdy, *antipod;
    const double  *I_bdy;                     // [NB*F]
    const double  *ray_w, *dshift_max;
    double         lstart_threshold;          // line[l] * this < shift_min * nu_f => skipped everywhere
    double        *J, *Lam;                   // [N*F]
    int32_t       *err_flag;
};

// get_eta_and_chi (simulation.tpp:385-481) for one lane: the line window is
// decided in FP64 exactly as the reference (skip while freq_diff > 11 width,
// include while freq_diff >= -11 width, the last line always reachable);
// profiles and sums in `real`, scaled by CHI_SCALE (and I_SCALE for eta).
template <typename real>
__device__ __forceinline__ void eta_chi(const RFArgs<real> &a, int32_t c, double fs, int lstart,
                                        real &eta, real &chi)
{
    const int L = a.L;
    const double *__restrict__ W = a.W11 + (int64_t)c * L;
    const real *__restrict__ IW = a.IW + (int64_t)c * L;
    const real *__restrict__ E = a.E + (int64_t)c * L;
    const real *__restrict__ K = a.K + (int64_t)c * L;
    int l = lstart;
    double fd, w11;
    {
#pragma clang fp contract(off)
        fd = fs - a.tb->line[l];
    }
    w11 = W[l];
    while (fd > w11) {
        if (l >= L - 1) break;
        l++;
        {
#pragma clang fp contract(off)
            fd = fs - a.tb->line[l];
        }
        w11 = W[l];
    }
    real e_acc = (real)0, c_acc = (real)(CHI_FLOOR * CHI_SCALE);
    const real nus = (real)fs;
    while (fd >= -w11) {
        const real x = (real)fd * IW[l];
        const real g = nus * Prec<real>::expneg(x * x);
        e_acc += E[l] * g;
        c_acc += K[l] * g;
        if (l >= L - 1) break;
        l++;
        {
#pragma clang fp contract(off)
            fd = fs - a.tb->line[l];
        }
        w11 = W[l];
    }
    eta = e_acc;
    chi = c_acc;
}

// Radiation::rescale_I_bdy, scalar branch (radiation.tpp:341-373): search the
// boundary cell's own frequency grid for the first bin >= freq (the linear
// search_with_notch from 0 is a lower bound), edge bins taken as they are,
// otherwise linear interpolation between bins notch-1 and notch.
__device__ __forceinline__ double rescale_I_bdy(const double *__restrict__ nup,
                                                const double *__restrict__ Ib, int F, double fs)
{
#pragma clang fp contract(off)
    int lo = 0, hi = F;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (nup[mid] < fs) lo = mid + 1; else hi = mid;
    }
    int notch = lo < F ? lo : F - 1;
    if (notch == 0 || notch == F - 1) return Ib[notch];
    const double nu1 = nup[notch], nu2 = nup[notch - 1];
    const double I1 = Ib[notch], I2 = Ib[notch - 1];
    return (I2 - I1) / (nu2 - nu1) * (fs - nu1) + I1;
}

// One ray of a pair, eliminated from its boundary end inwards up to the
// origin.  Returns f = F/(1+F) and Z of the point next to the origin and the
// optical depth D between that point and the origin (plus 1/D).
template <typename real>
struct SideState {
    real f, Z, D, invD;
};

template <typename real>
__device__ __forceinline__ SideState<real> sweep_side(const RFArgs<real> &a, int32_t o, int64_t q,
                                                      double nu_f, int lstart, real chi_o,
                                                      real S_o, double dsm, int F,
                                                      int32_t &err)
{
    const int64_t s0 = a.seg_off[q];
    const int ns = a.nseg[q];
    bool have_q = false, first_row = true;
    real chi_q = 0, S_q = 0, dZq = 0, Ib = 0;
    real f_prev = 0, Z_prev = 0, D_prev = 0, invD_prev = 0;

    auto step = [&](real chi, real S, real dZs) {
        if (have_q) {
            const real D = (real)0.5 * (chi_q + chi) * dZq;
            real Fn, Zn;
            if (first_row) {
                Fn = D * ((real)1 + (real)0.5 * D);
                const real Dn = Prec<real>::rcp((real)1 + Fn);
                Zn = (S_q * ((real)0.5 * D * D) + Ib * D) * Dn;
                f_prev = Fn * Dn;
                first_row = false;
            } else {
                const real rr = D * invD_prev;
                const real cc = (real)0.5 * (D_prev + D) * D;
                Fn = cc + rr * f_prev;
                const real Dn = Prec<real>::rcp((real)1 + Fn);
                Zn = (S_q * cc + rr * Z_prev) * Dn;
                f_prev = Fn * Dn;
            }
            Z_prev = Zn;
            D_prev = D;
            invD_prev = Prec<real>::rcp(D);
        }
        chi_q = chi;
        S_q = S;
        dZq = dZs;
        have_q = true;
    };

    auto point = [&](int32_t c, double sh, double dz) {
        double fs;
        {
#pragma clang fp contract(off)
            fs = sh * nu_f;
        }
        real eta, chi;
        eta_chi(a, c, fs, lstart, eta, chi);
        const real S = eta * Prec<real>::rcp(chi);
        if (!have_q) {
            // far end of the ray: incoming boundary intensity
            const int32_t b = a.cell2bdy[c];
            if (b < 0) {
                err |= ERR_NOT_BOUNDARY;
                Ib = (real)0;
            } else {
                Ib = (real)(rescale_I_bdy(a.nu + (int64_t)c * F, a.I_bdy + (int64_t)b * F, F, fs) * I_SCALE);
            }
        }
        step(chi, S, (real)(dz * (1.0 / CHI_SCALE)));
    };

    for (int j = ns - 1; j >= 0; --j) {
        const int32_t cj = a.seg_cell[s0 + j];
        const double sj = a.seg_shift[s0 + j];
        const double zj = a.seg_dZ[s0 + j];
        int32_t crt;
        double s_crt;
        if (j > 0) {
            crt = a.seg_cell[s0 + j - 1];
            s_crt = a.seg_shift[s0 + j - 1];
        } else {
            crt = o;
            s_crt = 1.0;
        }
        const StepSplit sp = split_step(s_crt, sj, zj, dsm);
        if (!sp.interp) {
            point(cj, sj, zj);
        } else {
            for (int m = sp.n; m >= 1; --m)
                point((m < sp.half) ? crt : cj, sub_shift(s_crt, m, sp.dshift), sp.dZ);
        }
    }
    // the origin closes the innermost row
    step(chi_o, S_o, (real)0);
    SideState<real> st;
    st.f = f_prev;
    st.Z = Z_prev;
    st.D = D_prev;
    st.invD = invD_prev;
    return st;
}

template <typename real>
__global__ __launch_bounds__(WAVE) void k_radiation_field(RFArgs<real> a)
{
    const int64_t wid = blockIdx.x;
    const int32_t o = (int32_t)(wid / a.nchunk);
    const int chunk = (int)(wid % a.nchunk);
    if (o >= a.N) return;
    const int F = a.F;
    const int f = chunk * WAVE + (int)threadIdx.x;
    const bool active = f < F;
    const int fa = active ? f : F - 1;
    const double nu_f = a.nu[(int64_t)o * F + fa];

    // first line that can be in the window anywhere on this cell's rays
    int lstart = 0;
    while (lstart < a.L - 1 && a.tb->line[lstart] * a.lstart_threshold < nu_f) lstart++;

    // the origin point: unshifted frequency (simulation.tpp:210-221)
    real eta_o, chi_o;
    eta_chi(a, o, nu_f, lstart, eta_o, chi_o);
    const real S_o = eta_o * Prec<real>::rcp(chi_o);
    const double dsm = a.dshift_max[o];
    const int32_t bo = a.cell2bdy[o];

    int32_t err = 0;
    double J_acc = 0.0, Lam_acc = 0.0;
    const int Rh = a.R / 2;
    for (int r = 0; r < Rh; r++) {
        const int ar = a.antipod[r];
        const double w2 = 2.0 * a.ray_w[r];
        const int64_t qr = (int64_t)o * a.R + r;
        const int64_t qa = (int64_t)o * a.R + ar;
        const int n_r = a.npts[qr];
        const int n_a = a.npts[qa];
        if (n_r == 0 && n_a == 0) {
            // ndep == 1: u = I_bdy of the origin's boundary entry (sim_radiation.tpp:243-264)
            if (bo < 0) { err |= ERR_EMPTY_SIDE; continue; }
            J_acc += w2 * a.I_bdy[(int64_t)bo * F + fa];
            continue;
        }
        real u, lam;
        if (n_r > 0 && n_a > 0) {
            const SideState<real> A = sweep_side(a, o, qa, nu_f, lstart, chi_o, S_o, dsm, F, err);
            const SideState<real> B = sweep_side(a, o, qr, nu_f, lstart, chi_o, S_o, dsm, F, err);
            const real h = (real)0.5 * (A.D + B.D);
            const real den = Prec<real>::rcp(h + A.f * A.invD + B.f * B.invD);
            u = (S_o * h + A.Z * A.invD + B.Z * B.invD) * den;
            lam = h * den;
        } else {
            const int64_t qs = (n_a > 0) ? qa : qr;
            const SideState<real> A = sweep_side(a, o, qs, nu_f, lstart, chi_o, S_o, dsm, F, err);
            real I_o = 0;
            if (bo < 0) err |= ERR_EMPTY_SIDE;
            else I_o = (real)(a.I_bdy[(int64_t)bo * F + fa] * I_SCALE);
            const real h2 = (real)0.5 * A.D * A.D;
            const real den = Prec<real>::rcp(h2 + A.D + A.f);
            u = (S_o * h2 + I_o * A.D + A.Z) * den;
            lam = h2 * den;
        }
        J_acc += w2 * ((double)u * (1.0 / I_SCALE));
        Lam_acc += w2 * (double)lam;
    }
    if (err) atomicOr(a.err_flag, err);
    if (active) {
        a.J[(int64_t)o * F + f] = J_acc;
        a.Lam[(int64_t)o * F + f] = Lam_acc * CHI_SCALE / (double)chi_o;
    }
}

template <typename real>
int launch_rf(mgr_ctx *ctx, const real *IW, const real *E, const real *K)
{
    RFArgs<real> a;
    a.N = ctx->N;
    a.R = (int)ctx->R;
    a.F = (int)ctx->F;
    a.L = (int)ctx->L;
    a.nchunk = (int)((ctx->F + WAVE - 1) / WAVE);
    a.tb = ctx->tables;
    a.nu = ctx->nu;
    a.W11 = ctx->W11;
    a.IW = IW;
    a.E = E;
    a.K = K;
    a.seg_off = ctx->seg_off;
    a.nseg = ctx->nseg;
    a.npts = ctx->npts;
    a.seg_cell = ctx->seg_cell;
    a.seg_shift = ctx->seg_shift;
    a.seg_dZ = ctx->seg_dZ;
    a.cell2bdy = ctx->cell2bdy;
    a.antipod = ctx->d_antipod;
    a.I_bdy = ctx->I_bdy;
    a.ray_w = ctx->ray_w;
    a.dshift_max = ctx->dshift_max;
    // line l is skipped at every point when line[l] (1 + 11 pw_max) (1 + 1e-9) < shift_min nu_f
    a.lstart_threshold = (1.0 + LINE_WINDOW * ctx->pw_max) * (1.0 + 1.0e-9) / ctx->shift_min;
    a.J = ctx->J;
    a.Lam = ctx->Lam;
    a.err_flag = ctx->err_flag;
    const int64_t nblocks = ctx->N * a.nchunk;
    hipLaunchKernelGGL(k_radiation_field<real>, dim3((unsigned)nblocks), dim3(WAVE), 0, ctx->stream, a);
    MGR_LAUNCH_CHECK();
    return MGR_OK;
}

}  // namespace

namespace mgr {

int launch_radiation_field(mgr_ctx *ctx, int precision)
{
    if (precision == MGR_FP64) return launch_rf<double>(ctx, ctx->IW64, ctx->E64, ctx->K64);
    return launch_rf<float>(ctx, ctx->IW32, ctx->E32, ctx->K32);
}

}  // namespace mgr

extern "C" int mgr_compute_radiation_field(mgr_ctx *ctx, int precision)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    if (precision != MGR_FP32 && precision != MGR_FP64) return ctx->fail(MGR_ERR_ARGUMENT, "precision must be MGR_FP32 or MGR_FP64");
    if (!ctx->spectral_set) return ctx->fail(MGR_ERR_STATE, "compute_spectral_discretisation first");
    if (!ctx->bdy_set) return ctx->fail(MGR_ERR_STATE, "compute_boundary_intensities first");
    if (!ctx->pops_set) return ctx->fail(MGR_ERR_STATE, "level populations not set");
    MGR_HIP(hipSetDevice(ctx->device));
    if (!ctx->traced) {
        int rc = launch_trace(ctx);
        if (rc) return rc;
    }
    hipEvent_t e0, e1;
    MGR_HIP(hipEventCreate(&e0));
    MGR_HIP(hipEventCreate(&e1));
    MGR_HIP(hipMemsetAsync(ctx->err_flag, 0, sizeof(int32_t), ctx->stream));
    MGR_HIP(hipEventRecord(e0, ctx->stream));
    int rc = launch_radiation_field(ctx, precision);
    if (rc) { (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); return rc; }
    MGR_HIP(hipEventRecord(e1, ctx->stream));
    MGR_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    ctx->last_rf_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    int32_t errv = 0;
    MGR_HIP(hipMemcpy(&errv, ctx->err_flag, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (errv) return ctx->fail(MGR_ERR_TRACE, "a ray pair ended off the boundary or an interior origin has an empty ray");
    ctx->J_set = true;
    return MGR_OK;
}
