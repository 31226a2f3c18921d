// mgr_radiation.hip -- radiation field: line emissivity/opacity along the
// traced ray pairs and the second-order Feautrier solve, fused in one kernel.
//
// Replaces Simulation::compute_radiation_field (src/Simulation/sim_radiation.tpp:56-338)
// with its inner calls Simulation::setup / get_eta_and_chi (simulation.tpp:185-481),
// Radiation::rescale_I_bdy (radiation.tpp:341-373), RayPair::solve
// (raypair_solver.tpp:200-388, n_off_diag 0 or 1: the same diagonal operator)
// and RayPair::update_Lambda (raypair_lambda.tpp:94-149); with scattering it
// also keeps each pair's u (Radiation::u, sim_radiation.tpp:226-262).  The
// image kernel (Simulation::compute_and_write_image) reuses the same staging
// and elimination on a rest-frame trace.
//
// Work mapping (CDNA4, 64-wide wavefronts): one workgroup per origin cell,
// one wavefront per chunk of 64 frequency bins (lane = bin).  The workgroup
// loops over the nrays/2 ray pairs of its origin.  For every ray it stages the
// cached traced steps, chunk by chunk, into LDS together with the line rows of
// their cells (11 width, 1/width and the emissivity/opacity prefactors per
// line), so that the sweep reads uniform LDS broadcasts instead of dependent
// global gathers.  The interpolation split of every step (set_data,
// geometry.tpp:100-207) is precomputed at trace time (FP32) or computed once
// while staging (FP64).  J and the Lambda*
// diagonal accumulate in FP64 registers over rays and are written once per
// (cell, bin): no atomics.  Workgroups are mapped to XCDs in contiguous origin
// ranges so that neighbouring origins share an L2.
//
// The tridiagonal Feautrier system of a ray pair is solved two-sided: each ray
// of the pair is eliminated from its boundary inwards in the Rybicki-Hummer
// (1+F) form and both eliminations meet at the origin, which yields u and the
// diagonal of the inverse there without storing or back-substituting the pair.
// With a = Dtau_{n-1}, b = Dtau_n the reference's F and Su recurrences, divided
// through by C_n, read F_n = (a+b) b/2 + (b/a) F_{n-1}/(1+F_{n-1}) and
// Z_n = (S_n (a+b) b/2 + (b/a) Z_{n-1}) / (1+F_n).  In the variables
// phi_n = F_n / b and zeta_n = Z_n / b every 1/a cancels:
//   h = (a+b)/2,  phi_n = h + phi_{n-1} q_{n-1},  q_n = 1/(1 + b phi_n),
//   zeta_n = (S_n h + zeta_{n-1}) q_n,
// one reciprocal per point.  The virtual start phi = q = 1, zeta = I_bdy, a = 0
// reproduces the boundary row F_0 = b(1 + b/2), Z_0 = (S b^2/2 + I b)/(1+F_0).
// At the origin, with h_o = (b_A + b_B)/2,
//   u = (S h_o + zeta_A + zeta_B) / (h_o + phi_A q_A + phi_B q_B),
//   Lambda = h_o / (h_o + phi_A q_A + phi_B q_B),
// and an empty side (origin on the boundary, incoming I_bdy of the origin's
// boundary cell) enters as the virtual start: phi q = 1, zeta = I, b = 0.
// phi stays below 1 + tau and q in (0, 1], so the FP32 variant neither
// overflows nor loses the optically thin limit.
// The kernels carry these in scaled variables, c = b/sqrt2, x = sqrt2 phi,
// z = sqrt2 zeta, y = x q + c and p = S c_prev + z_prev, in which
//   x_n = c_n + y_{n-1},  q_n = 1/(1 + c_n x_n),  y_n = x_n q_n + c_n,
//   z_n = (S_n c_n + p_n) q_n,
// so that no constant factor and no value from two points back is needed
// (no register copies around the loop); the virtual start is y = sqrt2,
// z = p = sqrt2 I, c = 0, and at the origin
//   u = (p_A + p_B) / (y_A + y_B),  Lambda = (c_A + c_B) / (y_A + y_B).
//
// Line profiles.  FP64 variant: the window of get_eta_and_chi is decided in
// FP64 exactly as the reference.  FP32 variant: each lane evaluates the lines
// whose +-11-width window can contain its bin at any point of the model (its
// candidate lines; lines outside contribute exactly zero in FP64 too, and
// inside the window but beyond ~10 widths exp(-x^2) underflows FP32), with the
// offset x = ((nu_f - nu_l) + delta nu_f) / width formed from the per-lane
// difference nu_f - nu_l and the point's shift delta = shift - 1, which keeps
// the profile argument accurate in FP32.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "mgr_common.hpp"
#include "mgr_ray.hpp"

using namespace mgr;

namespace {

constexpr int32_t ERR_NOT_BOUNDARY = 2;   // ray ends on a non-boundary cell
constexpr int32_t ERR_EMPTY_SIDE   = 4;   // empty ray at a non-boundary origin
constexpr int CH_MAX = 64;                // traced steps staged per chunk (at most)
constexpr size_t LDS_LIMIT = 64 * 1024;   // dynamic LDS per workgroup
constexpr int MAX_WAVES = 4;              // frequency chunks per workgroup

template <typename real> struct Prec;
template <> struct Prec<float> {
    using row_t = float4;   // (1/width, E, K, 0)
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
};
template <> struct Prec<double> {
    using row_t = double4;  // (11 width, 1/width, E, K)
    // the hardware reciprocal and two Newton steps (within an ulp or so of
    // 1/x; the elimination's arguments are normal numbers, chi scaled and
    // 1 + c x >= 1)
    static __device__ __forceinline__ double rcp(double x)
    {
        double r = __builtin_amdgcn_rcp(x);
        double e = __builtin_fma(-x, r, 1.0);
        r = __builtin_fma(r, e, r);
        e = __builtin_fma(-x, r, 1.0);
        return __builtin_fma(r, e, r);
    }
};

// One traced step in LDS.  Sub-point m = 1..n of the step lies at
// shift = s0 + m ds (FP64: s0 = shift of the current point, FP32: s0 and ds
// are shift - 1 and its increment) with path-length factor dz = dZ HDZ_SCALE and
// belongs to the current cell for m < half, to `cell` otherwise.
template <typename real>
struct alignas(32) Step {
    real    s0, ds, dz;   // first sub-point shift (offset), shift increment, dz'
    int32_t nh;           // n | half << 14 | out-of-window bits << 27 (STEP_*): one 16-byte read with the above
    int32_t cell;
};

// Steps staged for a kernel: the FP32 split of the radiation field, or the
// FP64 split (FP64 kernels, and the FP32 image, whose shifts stay FP64).
template <typename real, bool X64> struct StepOf { using type = Step<real>; };
template <> struct StepOf<float, true> { using type = Step<double>; };
template <typename real, bool X64> using step_t = typename StepOf<real, X64>::type;

// The sweep's view of staged step e: one 16-byte LDS read for an FP32 record.
__device__ __forceinline__ void read_step(const Step<float> *st, int e, float &s0, float &ds, float &dz, int &nh)
{
    const float4 v = reinterpret_cast<const float4 *>(st + e)[0];
    s0 = v.x;
    ds = v.y;
    dz = v.z;
    nh = __float_as_int(v.w);
}
__device__ __forceinline__ void read_step(const Step<double> *st, int e, double &s0, double &ds, double &dz, int &nh)
{
    s0 = st[e].s0;
    ds = st[e].ds;
    dz = st[e].dz;
    nh = st[e].nh;
}

template <typename real>
struct RFArgs {
    int64_t N;
    int     F, L, nchunk, nwg, ngroups, grid_real, ch;
    int     pair_stride;                      // every pair_stride-th ray pair, weights times pair_stride
    int32_t o_begin;                          // first origin of this launch
    size_t  row_offset;                       // byte offset of the row region in LDS
    const LineTables *tb;
    const double  *nu;                        // [N*F]
    const typename Prec<real>::row_t *rows;   // [N*L]
    const int64_t *seg_off;
    const int32_t *nseg, *npts, *seg_cell;
    const double  *seg_shift, *seg_dZ;
    const int4    *seg_f32;                   // FP32 split of every step (cell, n | half << 14, s0, ds)
    const float   *seg_hdz32;                 // its dZ * HDZ_SCALE
    const int32_t *far_cell;                  // last emitted point of each (origin, ray), -1: none
    const double  *far_shift;
    const float   *far32;                     // cached far-end intensities [q F + f] of the FP32 kernel (or null)
    const double  *far64;                     // the same for the FP64 kernel (or null)
    uint64_t       L_magic;                   // ceil(2^32 / L): idx / L = (idx * L_magic) >> 32
    const double  *origin_shift;              // FP64 shift of the origin per entry (null: 1, co-moving)
    const uint8_t *step_mask;                 // per step, out-of-window bit per 64-bin chunk (or null)
    float          noop_lo, noop_hi;          // an out-of-window sub-point after another is an exact no-op on
                                              // (y, z, p) when noop_lo < y < noop_hi (noop_lo = inf: never)
    bool           root_major;                // lane order (lane_bin)
    const int32_t *cell2bdy, *antipod;        // antipod: global ray indices
    const int32_t *cell_model;                // [N] model of each cell
    const int64_t *cell_qbase;                // [N] first (origin, ray) entry of the cell
    const int32_t *m_R, *m_rbase, *m_active;  // per model: rays, first global ray, still iterating
    const double  *I_bdy;                     // [NB*F]
    const double  *ray_w, *dshift_max;
    double         cand_lo, cand_hi;          // candidate-line bounds (FP32 variant)
    double         lstart_threshold;          // first line that can be in a window (FP64 variant)
    double        *J, *Lam;                   // [N*F]
    // RayPair::update_Lambda's sums: back-map of bin f to transition f_k[f]
    // (L: none) and root f_z[f], with the line data and profile widths
    const int32_t *f_k, *f_z;
    const double  *qw, *A_k, *freq_k, *pw;
    double        *Lpart;                     // [N][nchunk][L] one sum per wave and transition
    double        *u_out;                     // [R/2][N*F] per-pair u (scattering), or null
    int32_t       *err_flag;
    int            nband;                     // Lambda* bands (n_off_diag - 1) of k_lambda_band
    int64_t        band_S;                    // band slots per origin
    double        *band_part;                 // [N][nchunk][band_S][L]
    size_t         line_offset;               // FP64: byte offset of the line-frequency copy in LDS
};

// ---- line profiles ----------------------------------------------------------

// FP64: get_eta_and_chi (simulation.tpp:385-481) for one lane with the exact
// window: skip while freq_diff > 11 width, include while freq_diff >= -11 width.
__device__ __forceinline__ void eta_chi64(const RFArgs<double> &a, const double4 *__restrict__ row,
                                          const double *__restrict__ line, double fs, int lstart, double &eta,
                                          double &chi)
{
#pragma clang fp contract(off)
    const int L = a.L;
    int l = lstart;
    double fd = fs - line[l];
    double4 v = row[l];
    while (fd > v.x) {
        if (l >= L - 1) break;
        l++;
        fd = fs - line[l];
        v = row[l];
    }
    double e_acc = 0.0, c_acc = CHI_FLOOR * CHI_SCALE;
    while (fd >= -v.x) {
        const double x = fd * v.y;
        const double g = fs * exp(-(x * x));
        e_acc += v.z * g;
        c_acc += v.w * g;
        if (l >= L - 1) break;
        l++;
        fd = fs - line[l];
        v = row[l];
    }
    eta = e_acc;
    chi = c_acc;
}

// FP32: the lane's candidate lines la .. la+nc-1.  The first candidate is
// evaluated unconditionally (a lane without candidates carries an offset so
// large that its term is exactly zero); further candidates behind
// wave-uniform flags.  Rows hold 1/width scaled by sqrt(log2 e), so
// exp(-x^2) = exp2(-x'^2).
struct Lane32 {
    float  nuf;        // nu_f
    float  A0, A1;     // nu_f - nu_l of the first two candidates (or 1e30)
    int    off0, off1; // byte offsets of their rows
    int    la, nc;
    bool   has2, has3; // wave-uniform: some lane has 2 / more than 2 candidates
    double nu_f64;
    double A0d, A1d;   // the same offsets in FP64 (FP32 image variant)
};

__device__ __forceinline__ float rfl(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x)));
}

__device__ __forceinline__ void eta_chi32(const RFArgs<float> &a, const float4 *__restrict__ row,
                                          const Lane32 &ln, float delta, float &eta, float &chi)
{
#pragma clang fp contract(off)
    const char *base = reinterpret_cast<const char *>(row);
    // nu_f (1 + delta) and (nu_f delta + nu_f - nu_l), each with one rounding
    const float nus = fmaf(delta, ln.nuf, ln.nuf);
    const float4 v0 = *reinterpret_cast<const float4 *>(base + ln.off0);
    const float x0 = fmaf(delta, ln.nuf, ln.A0) * v0.x;
    const float g0 = nus * __builtin_amdgcn_exp2f(-(x0 * x0));
    float e_acc = v0.y * g0;
    float c_acc = fmaf(v0.z, g0, (float)(CHI_FLOOR * CHI_SCALE));
    if (ln.has2) {
        const float4 v1 = *reinterpret_cast<const float4 *>(base + ln.off1);
        const float x1 = fmaf(delta, ln.nuf, ln.A1) * v1.x;
        const float g1 = nus * __builtin_amdgcn_exp2f(-(x1 * x1));
        e_acc = fmaf(v1.y, g1, e_acc);
        c_acc = fmaf(v1.z, g1, c_acc);
        if (ln.has3) {
            for (int c = 2; c < ln.nc; c++) {
                const int l = ln.la + c;
                const float4 v = row[l];
                const float x = fmaf(delta, ln.nuf, (float)(ln.nu_f64 - a.tb->line[l])) * v.x;
                const float g = nus * __builtin_amdgcn_exp2f(-(x * x));
                e_acc = fmaf(v.y, g, e_acc);
                c_acc = fmaf(v.z, g, c_acc);
            }
        }
    }
    eta = e_acc;
    chi = c_acc;
}

// FP32 image variant: the shift s (absolute, FP64) and the offset from each
// candidate line in FP64, the profile and the sums in FP32.  An image bin can
// sit far out in a line's wing and meet the line only where the rest-frame
// shift cancels its offset; both terms are ~1e-4 nu there, and their FP32
// rounding (6e-12 nu) would be a visible fraction of a line width.
__device__ __forceinline__ void eta_chi32x(const RFArgs<float> &a, const float4 *__restrict__ row,
                                           const Lane32 &ln, double s, float &eta, float &chi)
{
    const char *base = reinterpret_cast<const char *>(row);
    const double t = (s - 1.0) * ln.nu_f64;
    const float nus = (float)(t + ln.nu_f64);
    const float4 v0 = *reinterpret_cast<const float4 *>(base + ln.off0);
    const float x0 = (float)(t + ln.A0d) * v0.x;
    const float g0 = nus * __builtin_amdgcn_exp2f(-(x0 * x0));
    float e_acc = v0.y * g0;
    float c_acc = fmaf(v0.z, g0, (float)(CHI_FLOOR * CHI_SCALE));
    if (ln.has2) {
        const float4 v1 = *reinterpret_cast<const float4 *>(base + ln.off1);
        const float x1 = (float)(t + ln.A1d) * v1.x;
        const float g1 = nus * __builtin_amdgcn_exp2f(-(x1 * x1));
        e_acc = fmaf(v1.y, g1, e_acc);
        c_acc = fmaf(v1.z, g1, c_acc);
        if (ln.has3) {
            for (int c = 2; c < ln.nc; c++) {
                const int l = ln.la + c;
                const float4 v = row[l];
                const float x = (float)(t + (ln.nu_f64 - a.tb->line[l])) * v.x;
                const float g = nus * __builtin_amdgcn_exp2f(-(x * x));
                e_acc = fmaf(v.y, g, e_acc);
                c_acc = fmaf(v.z, g, c_acc);
            }
        }
    }
    eta = e_acc;
    chi = c_acc;
}

// Radiation::rescale_I_bdy, scalar branch (radiation.tpp:341-373), at the
// count lo of the boundary cell's bins below the frequency (the grid is
// ascending, so that is the first bin >= freq, where search_with_notch from
// 0 stops): edge bins as they are, otherwise linear interpolation.
__device__ __forceinline__ double rescale_at(const double *__restrict__ nup, const double *__restrict__ Ib,
                                             int F, double fs, int lo)
{
#pragma clang fp contract(off)
    const int notch = lo < F ? lo : F - 1;
    if (notch == 0 || notch == F - 1) return Ib[notch];
    const double nu1 = nup[notch], nu2 = nup[notch - 1];
    const double I1 = Ib[notch], I2 = Ib[notch - 1];
    return (I2 - I1) / (nu2 - nu1) * (fs - nu1) + I1;
}

// The largest power of two <= F: the first step of the searches below.
__device__ __forceinline__ int search_top(int F) { return 1 << (31 - __builtin_clz((unsigned)F)); }

// Radiation::rescale_I_bdy of the boundary cell whose bins are nup (its
// intensities Ib) at frequency fs.  The count of bins below fs is found by
// halving steps from search_top(F): the same probes for every lane.
__device__ __forceinline__ double rescale_I_bdy(const double *__restrict__ nup,
                                                const double *__restrict__ Ib, int F, double fs)
{
    int lo = 0;
    for (int step = search_top(F); step > 0; step >>= 1)
        if (lo + step <= F && nup[lo + step - 1] < fs) lo += step;
    return rescale_at(nup, Ib, F, fs, lo);
}

// ---- Feautrier elimination state of one side ----------------------------------

__device__ __forceinline__ float fmad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }

// The recurrences are written with explicit fused multiply-adds (and
// contraction off), so that their rounding does not depend on how the
// compiler schedules or unrolls the loops around them.
template <typename real>
struct Sweep {
    real S_q, chi_q, hdz_q, cq;   // pending (outer) point: S, chi, its dz' and chi dz'
    real c, y, z, p;              // last closed interval and the scaled state

    // Virtual start.  With hdz_q = cq = 0 a step closes nothing: c = 0,
    // x = y, q = 1 exactly, so the first point may be stepped like any other.
    __device__ __forceinline__ void init(real I_far)
    {
        S_q = chi_q = hdz_q = cq = 0;
        c = 0;
        y = (real)SQRT2;
        z = (real)SQRT2 * I_far;
        p = z;
    }

    __device__ __forceinline__ void pend(real S, real chi, real hdz)
    {
#pragma clang fp contract(off)
        p = fmad(S, c, z);
        S_q = S;
        chi_q = chi;
        hdz_q = hdz;
        cq = chi * hdz;
    }

    // close the pending row with the next point (S, chi) inwards; the
    // interval takes the pending point's dz'
    __device__ __forceinline__ void step(real chi)
    {
#pragma clang fp contract(off)
        c = fmad(chi, hdz_q, cq);
        close_row();
    }

    // outward sweeps: the interval takes the new (outer) point's dz'
    __device__ __forceinline__ void step_out(real chi, real hdz)
    {
#pragma clang fp contract(off)
        c = (chi_q + chi) * hdz;
        close_row();
    }

    __device__ __forceinline__ void close_row()
    {
#pragma clang fp contract(off)
        const real x = c + y;
        const real q = Prec<real>::rcp(fmad(c, x, (real)1));
        y = fmad(x, q, c);
        z = fmad(S_q, c, p) * q;
    }
};

// ---- staging ----------------------------------------------------------------

// Stage the entries e = 0..cnt-1 of chunk [a, b): entry e holds step j = a-1+e
// (j = -1 is the origin) and the line rows of its cell.
template <typename real, bool X64 = false>
__device__ __forceinline__ void stage_chunk(const RFArgs<real> &a, int32_t o, int64_t q, int j0,
                                            int cnt, double dsm, step_t<real, X64> *__restrict__ st,
                                            typename Prec<real>::row_t *__restrict__ rw)
{
    const int64_t s0 = a.seg_off[q];
    const int nthr = blockDim.x;
    for (int e = threadIdx.x; e < cnt; e += nthr) {
        const int j = j0 + e;
        step_t<real, X64> s;
        if (j < 0) {
            s.cell = o; s.nh = 0; s.s0 = 0; s.ds = 0; s.dz = 0;
        } else if constexpr (sizeof(real) == 4 && !X64) {
            // split at trace time with the same FP64 arithmetic as below
            const int4 v = a.seg_f32[s0 + j];
            s.cell = v.x;
            s.nh = v.y | (a.step_mask ? (int)a.step_mask[s0 + j] << STEP_OUT_SHIFT : 0);
            s.s0 = __int_as_float(v.z);
            s.ds = __int_as_float(v.w);
            s.dz = a.seg_hdz32[s0 + j];
        } else {
            const int32_t c = a.seg_cell[s0 + j];
            const double sn = a.seg_shift[s0 + j];
            const double z = a.seg_dZ[s0 + j];
            const double sc = (j > 0) ? a.seg_shift[s0 + j - 1] : (a.origin_shift ? a.origin_shift[q] : 1.0);
            const StepSplit sp = split_step(sc, sn, z, dsm);
            s.cell = c;
            s.nh = sp.n | ((sp.interp ? sp.half : 0) << STEP_HALF_SHIFT) |
                   (a.step_mask ? (int)a.step_mask[s0 + j] << STEP_OUT_SHIFT : 0);
            s.s0 = sp.interp ? sc : sn;               // absolute FP64 shift
            s.ds = sp.dshift;
            s.dz = sp.dZ * HDZ_SCALE;                 // dz'
        }
        st[e] = s;
    }
    __syncthreads();
    const int L = a.L;
    for (int idx = threadIdx.x; idx < cnt * L; idx += nthr) {
        // idx / L by a multiply with ceil(2^32 / L): exact for idx < 2^32 / L
        const int e = (int)(((uint64_t)(uint32_t)idx * a.L_magic) >> 32), l = idx - e * L;
        rw[idx] = a.rows[(int64_t)st[e].cell * L + l];
    }
    __syncthreads();
}

template <typename real>
struct LaneCtx;   // per-lane constants of the origin

template <>
struct LaneCtx<float> {
    Lane32 ln;
};
template <>
struct LaneCtx<double> {
    double nu_f;
    int    lstart;
    const double *line;   // line frequencies (a copy in LDS in the field kernel)
};

template <typename real>
__device__ __forceinline__ void eval_point(const RFArgs<real> &a, const LaneCtx<real> &lc,
                                           const typename Prec<real>::row_t *__restrict__ row,
                                           real s, real &eta, real &chi);

template <>
__device__ __forceinline__ void eval_point<float>(const RFArgs<float> &a, const LaneCtx<float> &lc,
                                                  const float4 *__restrict__ row, float s,
                                                  float &eta, float &chi)
{
    eta_chi32(a, row, lc.ln, s, eta, chi);
}

template <>
__device__ __forceinline__ void eval_point<double>(const RFArgs<double> &a, const LaneCtx<double> &lc,
                                                   const double4 *__restrict__ row, double s,
                                                   double &eta, double &chi)
{
    double fs;
    {
#pragma clang fp contract(off)
        fs = s * lc.nu_f;
    }
    eta_chi64(a, row, lc.line, fs, lc.lstart, eta, chi);
}

// The point evaluation of a kernel: eval_point, or the FP32 image variant
// with FP64 shifts.
template <typename real, bool X64, typename SR>
__device__ __forceinline__ void eval_at(const RFArgs<real> &a, const LaneCtx<real> &lc,
                                        const typename Prec<real>::row_t *__restrict__ row, SR s,
                                        real &eta, real &chi)
{
    if constexpr (X64) eta_chi32x(a, row, lc.ln, s, eta, chi);
    else eval_point<real>(a, lc, row, s, eta, chi);
}

// shift of sub-point m: FP64 exactly s0 + m ds as Geometry::set_data;
// FP32 the offset shift - 1 in one fused multiply-add.
__device__ __forceinline__ double sub_point_shift(double s0, int m, double ds)
{
#pragma clang fp contract(off)
    return s0 + (double)m * ds;
}
__device__ __forceinline__ float sub_point_shift(float s0, int m, float ds)
{
    return fmaf((float)m, ds, s0);
}

__device__ __forceinline__ int rfl_int(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float rfl_real(float x) { return rfl(x); }
__device__ __forceinline__ double rfl_real(double x) { return x; }

// Incoming intensity at the far end of entry q (FP64, unscaled): the
// boundary intensity of the cell of the last emitted point, rescaled to that
// point's frequency shift * nu_f (Simulation::setup, simulation.tpp:266-272,
// 327-333); 0 for a ray without points.
template <typename real>
__device__ __forceinline__ double far_intensity(const RFArgs<real> &a, int64_t q, double nu_f, int F,
                                                int32_t &err)
{
    const int32_t cf = a.far_cell[q];
    if (cf < 0) return 0.0;
    const int32_t b = a.cell2bdy[cf];
    if (b < 0) {
        err |= ERR_NOT_BOUNDARY;
        return 0.0;
    }
    double fs;
    {
#pragma clang fp contract(off)
        fs = a.far_shift[q] * nu_f;
    }
    return rescale_I_bdy(a.nu + (int64_t)cf * F, a.I_bdy + (int64_t)b * F, F, fs);
}

// Scaled incoming intensity of entry q in the sweep's precision: the cache of
// that precision when it is built (the same value, computed once per trace,
// frequency grid and boundary intensities), else far_intensity.
template <typename real>
__device__ __forceinline__ real far_in(const RFArgs<real> &a, int64_t q, double nu_f, int F, int f, int32_t &err)
{
    if constexpr (sizeof(real) == 4) {
        if (a.far32) return a.far32[q * F + f];
    } else {
        if (a.far64) return a.far64[q * F + f];
    }
    return (real)(far_intensity(a, q, nu_f, F, err) * I_SCALE);
}

// The far-end intensity cache: one wave per (origin, ray) entry, its lanes
// over the bins, each the value far_in computes in the kernel of precision T
// (far_intensity with the entry's far cell, boundary and shift read once).
template <typename T>
__global__ __launch_bounds__(256) void k_far_cache(RFArgs<float> a, int Rmax, T *__restrict__ out,
                                                   int32_t *__restrict__ err_flag)
{
    const int64_t e = (int64_t)blockIdx.x * (blockDim.x / WAVE) + rfl_int((int)(threadIdx.x / WAVE));
    if (e >= a.N * Rmax) return;
    const int32_t o = (int32_t)(e / Rmax);
    const int r = (int)(e - (int64_t)o * Rmax);
    if (r >= a.m_R[a.cell_model[o]]) return;
    const int64_t q = a.cell_qbase[o] + r;
    const int F = a.F;
    const int lane = (int)(threadIdx.x & (WAVE - 1));
    T *dst = out + q * F;
    const int32_t cf = a.far_cell[q];
    const int32_t b = cf < 0 ? -1 : a.cell2bdy[cf];
    if (b < 0) {   // no points (0), or a far cell off the boundary (0 and an error)
        if (cf >= 0 && lane == 0) atomicOr(err_flag, ERR_NOT_BOUNDARY);
        for (int f = lane; f < F; f += WAVE) dst[f] = (T)(0.0 * I_SCALE);
        return;
    }
    const double fsh = a.far_shift[q];
    const double *nup = a.nu + (int64_t)cf * F;
    const double *Ib = a.I_bdy + (int64_t)b * F;
    const double *nuo = a.nu + (int64_t)o * F;
    for (int f = lane; f < F; f += WAVE) {
        double fs;
        {
#pragma clang fp contract(off)
            fs = fsh * nuo[f];
        }
        dst[f] = (T)(rescale_I_bdy(nup, Ib, F, fs) * I_SCALE);
    }
}

// One ray of a pair, eliminated from its boundary end (incoming intensity
// I_far, scaled) inwards to the origin, which it leaves pending.
template <typename real, bool X64 = false>
__device__ __forceinline__ Sweep<real> sweep_side(const RFArgs<real> &a, const LaneCtx<real> &lc,
                                                  int32_t o, int64_t q, real I_far, real S_o,
                                                  real chi_o, double dsm,
                                                  step_t<real, X64> *__restrict__ st,
                                                  typename Prec<real>::row_t *__restrict__ rw,
                                                  int chunk = STEP_MASK_CHUNKS)
{
    using row_t = typename Prec<real>::row_t;
    Sweep<real> sw;
    sw.init(I_far);
    const real chi_floor = (real)(CHI_FLOOR * CHI_SCALE);
    // this wave's out-of-window bit in the staged nh word (none past the masked chunks)
    const unsigned cbit = chunk < STEP_MASK_CHUNKS ? 1u << (STEP_OUT_SHIFT + chunk) : 0u;
    const int ns = a.nseg[q];
    const int L = a.L;
    const int CH = a.ch;
    // Wave-uniform: the pending point is an out-of-window sub-point and y lies
    // in (noop_lo, noop_hi) on every lane.  Then each further out-of-window
    // sub-point closes its row with c <= 2 chi_floor dz'_max, below half an
    // ulp of y and with c y below 2^-25, so that x = y, q = 1, y stays y and,
    // with S = 0 pending and p = z, z and p stay as they are: only the pending
    // point's dz' (and chi dz') moves on.  The result is the same bits as
    // stepping every sub-point.
    bool noop = false;
    for (int b_ = ns; b_ > 0; b_ -= CH) {
        const int a_ = b_ > CH ? b_ - CH : 0;
        const int cnt = b_ - a_ + 1;
        __syncthreads();   // the previous chunk is consumed by every wave
        stage_chunk<real, X64>(a, o, q, a_ - 1, cnt, dsm, st, rw);
        int e = cnt - 1;
        while (e >= 1) {
            decltype(st->s0) s0, ds, dzs;
            int nhs;
            read_step(st, e, s0, ds, dzs, nhs);   // s0 and ds stay in VGPRs: one scalar operand per FMA
            int nh = rfl_int(nhs);                // one readfirstlane: n, half and the chunk bits
            if (!((unsigned)nh & cbit)) {
                noop = false;
                // a run of steps with points inside a line window, in a loop
                // of its own so that the sweep state keeps its registers
                for (;;) {
                    const int n = nh & STEP_N_MASK;
                    __builtin_assume(n >= 1);   // a traced step emits 1 .. 10000 points (0 fails the trace)
                    const int half = (nh >> STEP_HALF_SHIFT) & STEP_HALF_MASK;
                    const real hdz = (real)rfl_real(dzs);
                    // sub-points m >= half lie in the step's cell (row e), m < half in
                    // the previous one (row e-1); two loops keep each row's address
                    // loop-invariant
                    const row_t *rn = rw + e * L;
                    const row_t *rc = rw + (e - 1) * L;
                    const int mlo = half > 1 ? half : 1;
                    for (int m = n; m >= mlo; --m) {
                        real eta, chi;
                        eval_at<real, X64>(a, lc, rn, sub_point_shift(s0, m, ds), eta, chi);
                        const real S = eta * Prec<real>::rcp(chi);
                        sw.step(chi);   // on the first point an exact no-op (c = 0, q = 1)
                        sw.pend(S, chi, hdz);
                    }
                    for (int m = (half - 1 < n ? half - 1 : n); m >= 1; --m) {
                        real eta, chi;
                        eval_at<real, X64>(a, lc, rc, sub_point_shift(s0, m, ds), eta, chi);
                        const real S = eta * Prec<real>::rcp(chi);
                        sw.step(chi);
                        sw.pend(S, chi, hdz);
                    }
                    if (--e < 1) break;
                    read_step(st, e, s0, ds, dzs, nhs);
                    nh = rfl_int(nhs);
                    if ((unsigned)nh & cbit) break;
                }
                if (e < 1) break;
            }
            // step e: no bin of this wave meets a line window in it, so every
            // sub-point has eta = 0 and chi at the floor, as evaluated
            const int n = nh & STEP_N_MASK;
            __builtin_assume(n >= 1);
            const real hdz = (real)rfl_real(dzs);
            if (!noop) {
                // the first closes the row of the last evaluated point
                sw.step(chi_floor);
                sw.pend((real)0, chi_floor, hdz);
                if constexpr (sizeof(real) == 4)
                    noop = __all(sw.y > a.noop_lo && sw.y < a.noop_hi);
                if (!noop) {
                    for (int m = n - 1; m >= 1; --m) {
                        sw.step(chi_floor);
                        sw.pend((real)0, chi_floor, hdz);
                    }
                    --e;
                    continue;
                }
            }
            // a no-op stretch: pass over the following out-of-window steps of
            // the chunk, keeping the dz' of the last one
            int last = e;
            while (last > 1 && ((unsigned)rfl_int(st[last - 1].nh) & cbit)) last--;
            const real hdz_l = last == e ? hdz : (real)rfl_real(st[last].dz);
            sw.hdz_q = hdz_l;
            sw.cq = chi_floor * hdz_l;
            e = last - 1;
        }
    }
    sw.step(chi_o);   // the origin closes the innermost row and is left pending
    sw.pend(S_o, chi_o, (real)0);
    return sw;
}

// Continues an elimination outwards along entry q, from the origin (the
// pending point of sw) to the far end of the ray.  Each point closes the
// pending row with its own dZ (the outer point's, as Simulation::setup pairs
// them, simulation.tpp:306) and becomes the pending point; on return the
// far-end point is pending.
template <typename real, bool X64 = false>
__device__ __forceinline__ void sweep_outward(const RFArgs<real> &a, const LaneCtx<real> &lc, int32_t o,
                                              int64_t q, double dsm, Sweep<real> &sw,
                                              step_t<real, X64> *__restrict__ st,
                                              typename Prec<real>::row_t *__restrict__ rw)
{
    using row_t = typename Prec<real>::row_t;
    const int ns = a.nseg[q];
    const int L = a.L;
    const int CH = a.ch;
    for (int a_ = 0; a_ < ns; a_ += CH) {
        const int b_ = a_ + CH < ns ? a_ + CH : ns;
        const int cnt = b_ - a_ + 1;
        __syncthreads();   // the previous chunk is consumed by every wave
        stage_chunk<real, X64>(a, o, q, a_ - 1, cnt, dsm, st, rw);
        for (int e = 1; e < cnt; ++e) {
            const int nh = rfl_int(st[e].nh);   // one readfirstlane: n and half are wave-uniform
            const int n = nh & STEP_N_MASK;
            const int half = (nh >> STEP_HALF_SHIFT) & STEP_HALF_MASK;
            const auto s0 = st[e].s0;
            const auto ds = st[e].ds;
            const real hdz = (real)rfl_real(st[e].dz);
            const row_t *rn = rw + e * L;
            const row_t *rc = rw + (e - 1) * L;
            const int mhi = half - 1 < n ? half - 1 : n;
            for (int m = 1; m <= mhi; ++m) {
                real eta, chi;
                eval_at<real, X64>(a, lc, rc, sub_point_shift(s0, m, ds), eta, chi);
                const real S = eta * Prec<real>::rcp(chi);
                sw.step_out(chi, hdz);
                sw.pend(S, chi, hdz);
            }
            for (int m = half > 1 ? half : 1; m <= n; ++m) {
                real eta, chi;
                eval_at<real, X64>(a, lc, rn, sub_point_shift(s0, m, ds), eta, chi);
                const real S = eta * Prec<real>::rcp(chi);
                sw.step_out(chi, hdz);
                sw.pend(S, chi, hdz);
            }
        }
    }
}

// u at the pending end point of a full sweep, closed by the incoming
// intensity I (scaled) there: the boundary row of RayPair::solve's last point
// in the elimination variables.
template <typename real>
__device__ __forceinline__ real end_u(const Sweep<real> &sw, real I)
{
    return (sw.p + (real)SQRT2 * I) * Prec<real>::rcp(sw.y + (real)SQRT2);
}

// Per-lane constants of origin o's bin and its own emissivity and opacity at
// the unshifted frequency nu_f (Simulation::setup takes the origin at
// nu(o, f), simulation.tpp:209-222).  ONE: no bin has more than one candidate
// line (known on the host), so the second-line code is compiled out.
template <typename real, bool X64 = false, bool ONE = false>
__device__ __forceinline__ void origin_point(const RFArgs<real> &a, double nu_f,
                                             const typename Prec<real>::row_t *__restrict__ orow,
                                             LaneCtx<real> &lc, real &S_o, real &chi_o)
{
    real eta_o;
    if constexpr (sizeof(real) == 4) {
        // candidate lines: nu_l (1 - w) <= s_max nu_f and nu_l (1 + w) >= s_min nu_f
        int la = 0;
        while (la < a.L && a.tb->line[la] * a.cand_lo < nu_f) la++;
        int lb = la;
        while (lb < a.L && a.tb->line[lb] * a.cand_hi <= nu_f) lb++;
        const int nc = lb - la;
        const int l0 = la < a.L ? la : a.L - 1;
        const int l1 = la + 1 < a.L ? la + 1 : a.L - 1;
        lc.ln.la = l0;
        lc.ln.nc = nc;
        lc.ln.nuf = (float)nu_f;
        lc.ln.nu_f64 = nu_f;
        lc.ln.A0 = (nc > 0) ? (float)(nu_f - a.tb->line[l0]) : 1.0e30f;
        lc.ln.A1 = (nc > 1) ? (float)(nu_f - a.tb->line[l1]) : 1.0e30f;
        lc.ln.A0d = (nc > 0) ? nu_f - a.tb->line[l0] : 1.0e30;
        lc.ln.A1d = (nc > 1) ? nu_f - a.tb->line[l1] : 1.0e30;
        lc.ln.off0 = l0 * (int)sizeof(float4);
        lc.ln.off1 = l1 * (int)sizeof(float4);
        lc.ln.has2 = ONE ? false : __any(nc > 1);
        lc.ln.has3 = ONE ? false : __any(nc > 2);
        if constexpr (X64) eta_chi32x(a, orow, lc.ln, 1.0, eta_o, chi_o);
        else eval_point<real>(a, lc, orow, (real)0, eta_o, chi_o);
    } else {
        int ls = 0;
        while (ls < a.L - 1 && a.tb->line[ls] * a.lstart_threshold < nu_f) ls++;
        lc.nu_f = nu_f;
        lc.lstart = ls;
        eval_point<real>(a, lc, orow, (real)1, eta_o, chi_o);
    }
    S_o = eta_o * Prec<real>::rcp(chi_o);
}

// This wave's share of the Lambda* diagonal per transition
// (RayPair::update_Lambda, raypair_lambda.tpp:94-149, n_off_diag 0):
// Ls(p,k) = sum over the bins f of transition k of
// A_k w_z nu_f phi_k(nu_f) Lam(p,f), with Lam = sum_r 2 w_r L_diag / chi as
// the kernel accumulates it.  The wave sums its lanes' terms per transition
// in a fixed butterfly order and lane k writes transition k's sum;
// k_lambda_sum adds the waves of a cell in chunk order.
template <typename real>
__device__ __forceinline__ void lambda_partials(const RFArgs<real> &a, int32_t o, int chunk, bool active, int f,
                                                double nu_f, double lam_f)
{
    const int L = a.L;
    const int k = active ? a.f_k[f] : L;
    double term = 0.0;
    if (k < L) {
        const double freq_line = a.freq_k[k];
        const double iw = 1.0 / (freq_line * a.pw[(int64_t)a.tb->line_spec[k] * a.N + o]);   // the line's species
        const double x = iw * (nu_f - freq_line);
        const double prof = iw * INVERSE_SQRT_PI * (1.0 / exp(x * x));
        term = a.A_k[k] * a.qw[a.f_z[f]] * (nu_f * prof * lam_f);
    }
    const int lane = (int)(threadIdx.x & (WAVE - 1));
    // transitions present in this wave: [kmin, kmax]
    int kmin = k, kmax = k < L ? k : -1;
    for (int off = WAVE / 2; off >= 1; off >>= 1) {
        kmin = min(kmin, __shfl_xor(kmin, off));
        kmax = max(kmax, __shfl_xor(kmax, off));
    }
    double mine = 0.0;
    for (int kk = kmin; kk <= kmax; kk++) {
        double v = (k == kk) ? term : 0.0;
        for (int off = WAVE / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off);
        v = __shfl(v, 0);   // lane 0's order for every lane
        if (lane == kk) mine = v;
    }
    if (chunk < a.nchunk && lane < L)
        a.Lpart[((int64_t)o * a.nchunk + chunk) * L + lane] = mine;
}

// FP32: at least 8 waves per SIMD (at most 64 VGPRs), which hide the LDS
// latency of the line rows; FP64 is left to the compiler.
template <typename real, bool ONE = false>
__global__ __launch_bounds__(WAVE * MAX_WAVES) __attribute__((amdgpu_waves_per_eu(sizeof(real) == 4 ? 8 : 1)))
void k_radiation_field(RFArgs<real> a)
{
    using row_t = typename Prec<real>::row_t;
    extern __shared__ __align__(16) unsigned char smem[];
    Step<real> *st = reinterpret_cast<Step<real> *>(smem);
    row_t *rw = reinterpret_cast<row_t *>(smem + a.row_offset);

    // XCD-aware order: consecutive workgroups go to different XCDs, so hand
    // each XCD a contiguous range of origins.
    const int G = gridDim.x;
    const int per = G >> 3;
    const int bid = blockIdx.x;
    const int logical = (bid & 7) * per + (bid >> 3);
    if (logical >= a.grid_real) return;   // whole workgroup: before any barrier
    const int32_t o = a.o_begin + (int32_t)(logical / a.ngroups);
    const int32_t mdl = a.cell_model[o];
    if (!a.m_active[mdl]) return;         // converged model of a batch: whole workgroup
    const int group = logical - (o - a.o_begin) * a.ngroups;
    const int wave = threadIdx.x / WAVE;
    const int chunk = rfl_int(group * a.nwg + wave);   // wave-uniform: scalar skip tests
    const int F = a.F;
    const int slot = chunk * WAVE + (int)(threadIdx.x & (WAVE - 1));
    const bool active = slot < F;
    const int f = lane_bin(active ? slot : F - 1, a.L, F, a.root_major);
    const double nu_f = a.nu[(int64_t)o * F + f];
    const double dsm = a.dshift_max[o];
    const int32_t bo = a.cell2bdy[o];
    const row_t *orow = a.rows + (int64_t)o * a.L;

    LaneCtx<real> lc;
    if constexpr (sizeof(real) == 8) {
        // FP64: the line frequencies the profile windows read at every point, in LDS
        double *lt = reinterpret_cast<double *>(smem + a.line_offset);
        for (int i = (int)threadIdx.x; i < a.L; i += (int)blockDim.x) lt[i] = a.tb->line[i];
        __syncthreads();
        lc.line = lt;
    }
    real S_o, chi_o;
    origin_point<real, false, ONE>(a, nu_f, orow, lc, S_o, chi_o);

    int32_t err = 0;
    double J_acc = 0.0, Lam_acc = 0.0;
    const int rb = a.m_rbase[mdl];
    const int Rh = a.m_R[mdl] / 2;
    const int64_t qb = a.cell_qbase[o];
    const double wsc = 2.0 * (double)a.pair_stride;
    for (int r = 0; r < Rh; r += a.pair_stride) {
        const int ar = a.antipod[rb + r];
        const double w2 = wsc * a.ray_w[rb + r];
        const int64_t qr = qb + r;
        const int64_t qa = qb + (ar - rb);
        const int n_r = a.npts[qr];
        const int n_a = a.npts[qa];
        if (n_r == 0 && n_a == 0) {
            // ndep == 1: u = I_bdy of the origin's boundary entry (sim_radiation.tpp:243-264)
            if (bo < 0) { err |= ERR_EMPTY_SIDE; continue; }
            const double Ib = a.I_bdy[(int64_t)bo * F + f];
            J_acc += w2 * Ib;
            if (a.u_out && active) a.u_out[((int64_t)r * a.N + o) * F + f] = Ib;
            continue;
        }
        real u, lam;
        if (n_r > 0 && n_a > 0) {
            const real Ia = far_in<real>(a, qa, nu_f, F, f, err);
            const Sweep<real> A = sweep_side<real>(a, lc, o, qa, Ia, S_o, chi_o, dsm, st, rw, chunk);
            const real Ir = far_in<real>(a, qr, nu_f, F, f, err);
            const Sweep<real> B = sweep_side<real>(a, lc, o, qr, Ir, S_o, chi_o, dsm, st, rw, chunk);
            const real den = Prec<real>::rcp(A.y + B.y);
            u = (A.p + B.p) * den;
            lam = (A.c + B.c) * den;
        } else {
            const int64_t qs = (n_a > 0) ? qa : qr;
            const real Is = far_in<real>(a, qs, nu_f, F, f, err);
            const Sweep<real> A = sweep_side<real>(a, lc, o, qs, Is, S_o, chi_o, dsm, st, rw, chunk);
            real I_o = 0;
            if (bo < 0) err |= ERR_EMPTY_SIDE;
            else I_o = (real)(a.I_bdy[(int64_t)bo * F + f] * I_SCALE);
            const real den = Prec<real>::rcp(A.y + (real)SQRT2);
            u = (A.p + (real)SQRT2 * I_o) * den;
            lam = A.c * den;
        }
        const double u64 = (double)u * (1.0 / I_SCALE);
        J_acc += w2 * u64;
        Lam_acc += w2 * (double)lam;
        if (a.u_out && active) a.u_out[((int64_t)r * a.N + o) * F + f] = u64;   // radiation.u[R][ind]
    }
    if (err) atomicOr(a.err_flag, err);
    const double lam_f = Lam_acc * CHI_SCALE / (double)chi_o;
    if (active) {
        a.J[(int64_t)o * F + f] = J_acc;
        a.Lam[(int64_t)o * F + f] = lam_f;
    }
    lambda_partials(a, o, chunk, active, f, nu_f, lam_f);
}

// Image of ray r (Simulation::compute_and_write_image, sim_radiation.tpp:345-478;
// RayPair::solve_for_image, raypair_solver.tpp:523-579).  Entries 2o (the
// antipode, index 0 end of the pair) and 2o + 1 (ray r, index ndep-1 end) of
// the rest-frame trace.  The intensities leaving the pair are
// I_p = 2 u(ndep-1) - I_bdy_n and I_m = 2 u(0) - I_bdy_0 (RayPair::get_Ip_at_end,
// get_Im_at_front, raypair.tpp:212-229).  Instead of a stored forward
// elimination and back-substitution, each end's u comes from one full sweep
// that starts at the other end, passes the origin and closes at the end
// itself with its boundary row: two sweeps, no per-point storage.
template <typename real, bool X64>
__global__ __launch_bounds__(WAVE * MAX_WAVES) void k_image(RFArgs<real> a, double *__restrict__ I_m,
                                                            double *__restrict__ I_p)
{
    using row_t = typename Prec<real>::row_t;
    extern __shared__ __align__(16) unsigned char smem[];
    step_t<real, X64> *st = reinterpret_cast<step_t<real, X64> *>(smem);
    row_t *rw = reinterpret_cast<row_t *>(smem + a.row_offset);

    const int G = gridDim.x;
    const int per = G >> 3;
    const int bid = blockIdx.x;
    const int logical = (bid & 7) * per + (bid >> 3);
    if (logical >= a.grid_real) return;   // whole workgroup: before any barrier
    const int32_t o = a.o_begin + (int32_t)(logical / a.ngroups);
    const int group = logical - (o - a.o_begin) * a.ngroups;
    const int wave = threadIdx.x / WAVE;
    const int chunk = rfl_int(group * a.nwg + wave);   // wave-uniform: scalar skip tests
    const int F = a.F;
    const int f = chunk * WAVE + (int)(threadIdx.x & (WAVE - 1));
    const bool active = f < F;
    const int fa = active ? f : F - 1;
    const double nu_f = a.nu[(int64_t)o * F + fa];
    const double dsm = a.dshift_max[o];
    const int32_t bo = a.cell2bdy[o];
    const row_t *orow = a.rows + (int64_t)o * a.L;

    LaneCtx<real> lc;
    if constexpr (sizeof(real) == 8) lc.line = a.tb->line;
    real S_o, chi_o;
    origin_point<real, X64>(a, nu_f, orow, lc, S_o, chi_o);

    int32_t err = 0;
    const int64_t qa = 2 * (int64_t)o, qr = qa + 1;
    const int n_a = a.npts[qa];
    const int n_r = a.npts[qr];
    double Im = 0.0, Ip = 0.0;
    if (n_a == 0 && n_r == 0) {
        // too short a ray: both intensities are I_bdy of the origin's boundary entry
        if (bo < 0) err |= ERR_EMPTY_SIDE;
        else Im = Ip = a.I_bdy[(int64_t)bo * F + fa];
    } else {
        double I_o = 0.0;   // an empty side takes I_bdy of the origin's boundary entry
        if (n_a == 0 || n_r == 0) {
            if (bo < 0) err |= ERR_EMPTY_SIDE;
            else I_o = a.I_bdy[(int64_t)bo * F + fa];
        }
        const double I0 = n_a > 0 ? far_intensity(a, qa, nu_f, F, err) : I_o;
        const double In = n_r > 0 ? far_intensity(a, qr, nu_f, F, err) : I_o;
        const real I0s = (real)(I0 * I_SCALE);
        const real Ins = (real)(In * I_SCALE);
        real u_n, u_0;
        {
            // from the antipode's far end through the origin to ray r's end
            Sweep<real> A;
            if (n_a > 0) {
                A = sweep_side<real, X64>(a, lc, o, qa, I0s, S_o, chi_o, dsm, st, rw);
            } else {
                A.init(I0s);
                A.pend(S_o, chi_o, (real)0);
            }
            sweep_outward<real, X64>(a, lc, o, qr, dsm, A, st, rw);
            u_n = end_u(A, Ins);
        }
        {
            // and back
            Sweep<real> B;
            if (n_r > 0) {
                B = sweep_side<real, X64>(a, lc, o, qr, Ins, S_o, chi_o, dsm, st, rw);
            } else {
                B.init(Ins);
                B.pend(S_o, chi_o, (real)0);
            }
            sweep_outward<real, X64>(a, lc, o, qa, dsm, B, st, rw);
            u_0 = end_u(B, I0s);
        }
        Ip = 2.0 * ((double)u_n * (1.0 / I_SCALE)) - In;
        Im = 2.0 * ((double)u_0 * (1.0 / I_SCALE)) - I0;
    }
    if (err) atomicOr(a.err_flag, err);
    if (active) {
        I_m[(int64_t)o * F + f] = Im;
        I_p[(int64_t)o * F + f] = Ip;
    }
}

// ---- banded Lambda* (n_off_diag >= 2) ------------------------------------------

// Sub-points of traced step j of the entry at base: m = 1 .. n at shift
// s0 + m ds with dZ dz, in cell c_prev for m < half and c_next otherwise
// (Geometry::set_data, geometry.tpp:100-207).  step_raw loads what the split
// needs, split_points forms it, so that the loads of a step can be issued
// before the previous step's points are visited.
struct StepPoints {
    int     n, half;
    double  s0, ds, dz;
    int32_t c_prev, c_next;
};

struct StepRaw {
    double  sc, sn, z;
    int32_t c_prev, c_next;
};

__device__ __forceinline__ StepRaw step_raw(const RFArgs<double> &a, int64_t base, int j, int32_t o)
{
    StepRaw r;
    r.sn = a.seg_shift[base + j];
    r.sc = j > 0 ? a.seg_shift[base + j - 1] : 1.0;   // co-moving: the origin's shift is 1
    r.z = a.seg_dZ[base + j];
    r.c_prev = j > 0 ? a.seg_cell[base + j - 1] : o;
    r.c_next = a.seg_cell[base + j];
    return r;
}

__device__ __forceinline__ StepPoints split_points(const StepRaw &r, double dsm)
{
    StepPoints p;
    const StepSplit sp = split_step(r.sc, r.sn, r.z, dsm);
    p.n = sp.n;
    p.half = sp.interp ? sp.half : 0;
    p.s0 = sp.interp ? r.sc : r.sn;
    p.ds = sp.dshift;
    p.dz = sp.dZ;
    p.c_prev = r.c_prev;
    p.c_next = r.c_next;
    return p;
}

__device__ __forceinline__ StepPoints step_points(const RFArgs<double> &a, int64_t base, int j, int32_t o,
                                                  double dsm)
{
    return split_points(step_raw(a, base, j, o), dsm);
}

// The points of a ray pair in order: entry qx from its far end inwards, the
// origin, then entry qy outwards, at most ny points of it.  visit(t, cell,
// shift, dz, out) gets the visit count t and the dZ of the interval between
// the point and the one visited before it, which is the dZ of the interval's
// outer point (Simulation::setup, simulation.tpp:233-240, 292-299); out is
// set when the step's out-of-window bit `cbit` (of a.step_mask) is set: no
// line is then in the window of any bin of this wave (eta = 0, chi at the
// floor).
template <typename Visit>
__device__ __forceinline__ void walk_pair(const RFArgs<double> &a, int32_t o, int64_t qx, int64_t qy, int ny,
                                          double dsm, unsigned cbit, Visit &&visit)
{
    int t = 0;
    double dz_prev = 0.0;
    const int64_t bx = a.seg_off[qx];
    // each step's loads are issued before the previous step's points are visited
    int j = a.nseg[qx] - 1;
    StepRaw nx;
    uint8_t nmask = 0;
    if (j >= 0) {
        nx = step_raw(a, bx, j, o);
        nmask = cbit ? a.step_mask[bx + j] : 0;
    }
    for (; j >= 0; --j) {
        const StepPoints sp = split_points(nx, dsm);
        const bool out = (nmask & cbit) != 0;
        if (j > 0) {
            nx = step_raw(a, bx, j - 1, o);
            nmask = cbit ? a.step_mask[bx + j - 1] : 0;
        }
        for (int m = sp.n; m >= 1; --m) {
            visit(t++, m < sp.half ? sp.c_prev : sp.c_next, sub_point_shift(sp.s0, m, sp.ds), dz_prev, out);
            dz_prev = sp.dz;
        }
    }
    visit(t++, o, 1.0, dz_prev, false);
    const int64_t by = a.seg_off[qy];
    const int nsy = a.nseg[qy];
    for (int jy = 0; jy < nsy && ny > 0; ++jy) {
        const StepPoints sp = step_points(a, by, jy, o, dsm);
        const bool out = cbit && (a.step_mask[by + jy] & cbit);
        for (int m = 1; m <= sp.n && ny > 0; ++m, --ny)
            visit(t++, m < sp.half ? sp.c_prev : sp.c_next, sub_point_shift(sp.s0, m, sp.ds), sp.dz, out);
    }
}

// Thermodynamics::profile (thermodynamics.tpp:16-29) of width `width` at
// freq_diff, with the reference's scalar 1/exp.
__device__ __forceinline__ double profile64(double width, double freq_diff)
{
#pragma clang fp contract(off)
    const double iw = 1.0 / width;
    const double x = iw * freq_diff;
    return iw * INVERSE_SQRT_PI * (1.0 / exp(x * x));
}

// Per-lane window of the band kernel in LDS: array k, pair index n at
// ((k W) + n - (n_ar - nband)) WAVE + lane, W = 3 (nband + 1), covering
// n_ar - nband .. n_ar + 2 nband + 2.
constexpr int BW_F = 0, BW_IC = 1, BW_G = 2, BW_FP = 3, BW_CHI = 4, BW_ARRAYS = 5;

// The band elements of Lambda* (RayPair::update_Lambda, raypair_lambda.tpp:129-146,
// with get_L_lower / get_L_upper, raypair_lambda.tpp:44-84) in FP64.  One wave
// of 64 bins per (origin, chunk of bins).  Per ray pair the longer ray is the
// low side (n_ar = nA points, sim_radiation.tpp:164-171), and the reference's
// own variables are formed with its operation order (raypair_solver.tpp:200-388):
//   pass 1, from index 0 up: dtau, F and 1/C up to index n_ar + 2 nband + 1;
//   pass 2, from index last down: G down to index n_ar - nband + 1;
// then for band m = 1 .. nband
//   lower: L_lower[m][n_ar-m] = L_diag[n_ar-m] prod_{q=n_ar-m+1}^{n_ar+1} 1/(1+G_q),
//   upper: L_upper[m][n_ar+m] = L_diag[n_ar+2m+1] prod_{q=n_ar+2m}^{n_ar+m} 1/(1+F_q),
// each times A_k (w_z 2 w_r) nu phi(nu) / chi at its own point, where the
// pair's solve wrote it (the slots k_band_cells marks).  The wave sums its
// lanes per transition; k_band_sum adds the waves of an origin.
__global__ __launch_bounds__(WAVE) void k_lambda_band(RFArgs<double> a)
{
#pragma clang fp contract(off)
    extern __shared__ __align__(16) unsigned char smem[];
    double *win = reinterpret_cast<double *>(smem);
    const int G = gridDim.x;
    const int per = G >> 3;
    const int bid = blockIdx.x;
    const int logical = (bid & 7) * per + (bid >> 3);
    if (logical >= a.grid_real) return;
    const int32_t o = (int32_t)(logical / a.nchunk);
    const int chunk = logical - o * a.nchunk;
    const int lane = (int)threadIdx.x;
    const int F = a.F, L = a.L, nb = a.nband;
    const int W = 3 * (nb + 1);
    // the lane order of the field kernels (lane_bin), which the out-of-window bits follow
    const int slot = chunk * WAVE + lane;
    const bool active = slot < F;
    const int fa = lane_bin(active ? slot : F - 1, L, F, a.root_major);
    const int k = active ? a.f_k[fa] : L;
    const double nu_f = a.nu[(int64_t)o * F + fa];
    const double dsm = a.dshift_max[o];
    LaneCtx<double> lc;
    {
        // the line frequencies the profile windows read at every point, in LDS after the window arrays
        double *lt = win + (size_t)BW_ARRAYS * 3 * (nb + 1) * WAVE;
        for (int i = lane; i < L; i += WAVE) lt[i] = a.tb->line[i];
        __syncthreads();
        lc.line = lt;
    }
    double S_o, chi_o;
    origin_point<double>(a, nu_f, a.rows + (int64_t)o * L, lc, S_o, chi_o);
    const double freq_line = k < L ? a.freq_k[k] : 1.0;
    const double A_k = k < L ? a.A_k[k] : 0.0;
    const double w_z = k < L ? a.qw[a.f_z[fa]] : 0.0;
    auto wref = [&](int arr, int w) -> double & { return win[(arr * W + w) * WAVE + lane]; };
    const int64_t qb = a.cell_qbase[o];
    const int Rh = a.m_R[0] / 2;
    double *part = a.band_part + ((int64_t)o * a.nchunk + chunk) * a.band_S * L;
    // this wave's out-of-window bit (the FP64 kernel's masks), none past the masked chunks
    const unsigned cbit = (a.step_mask && chunk < STEP_MASK_CHUNKS) ? 1u << chunk : 0u;
    // chi of a point: only the opacity enters dtau; at an out-of-window point the floor
    auto opacity = [&](int32_t c, double s, bool out) -> double {
        double eta, chi = CHI_FLOOR * CHI_SCALE;
        if (!out) eval_point<double>(a, lc, a.rows + (int64_t)c * L, s, eta, chi);
        return chi * (1.0 / CHI_SCALE);
    };

    for (int r = 0; r < Rh; r++) {
        const int64_t qr = qb + r, qa = qb + a.antipod[r];
        const int n_r = a.npts[qr], n_a = a.npts[qa];
        const bool r_longer = n_r > n_a;
        const int64_t qA = r_longer ? qr : qa, qB = r_longer ? qa : qr;
        const int nA = r_longer ? n_r : n_a, nB = r_longer ? n_a : n_r;
        const int last = nA + nB;
        const int w0 = nA - nb;
        double dt0 = 0.0, dtL = 0.0;
        if (nB >= 1) {
            // pass 1: from index 0 (the far end of the longer ray) upwards
            double chi_p = 0.0, dt_p = 0.0, F_p = 0.0, i1F_p = 0.0;
            const int top = nA + 2 * nb + 2 < last ? nA + 2 * nb + 2 : last;
            walk_pair(a, o, qA, qB, top - nA, dsm, cbit, [&](int n, int32_t c, double s, double dz, bool out) {
                const double chi = opacity(c, s, out);
                const int w = n - w0;
                if (w >= 0 && w < W) {
                    // nu phi(nu) at the point's own cell and frequency, and chi
                    const double fs = s * nu_f;
                    wref(BW_FP, w) = k < L ? fs * profile64(freq_line * a.pw[c], fs - freq_line) : 0.0;
                    wref(BW_CHI, w) = chi;
                }
                if (n >= 1) {
                    const double dt = 0.5 * (chi_p + chi) * dz;   // dtau[n-1]
                    if (n == 1) {
                        dt0 = dt;
                        const double B0_min_C0 = 1.0 + 2.0 * (1.0 / dt);
                        F_p = 0.5 * B0_min_C0 * dt * dt;
                        i1F_p = 1.0 / (1.0 + F_p);
                        if (w - 1 >= 0 && w - 1 < W) wref(BW_F, w - 1) = F_p;
                    } else {
                        const double iA = 0.5 * (dt_p + dt) * dt_p;   // inverse_A[n-1]
                        const double iC = 0.5 * (dt_p + dt) * dt;     // inverse_C[n-1]
                        F_p = (1.0 + Prec<double>::rcp(iA) * F_p * i1F_p) * iC;  // F[n-1]
                        i1F_p = Prec<double>::rcp(1.0 + F_p);
                        if (w - 1 >= 0 && w - 1 < W) {
                            wref(BW_F, w - 1) = F_p;
                            wref(BW_IC, w - 1) = iC;
                        }
                    }
                    dt_p = dt;
                }
                chi_p = chi;
            });
            // pass 2: from index last (the far end of the shorter ray) downwards
            double G_p = 0.0, i1G_p = 0.0;
            chi_p = 0.0;
            dt_p = 0.0;
            const int bottom = nA - nb > 0 ? nA - nb : 0;
            walk_pair(a, o, qB, qA, nA - bottom, dsm, cbit, [&](int t, int32_t c, double s, double dz, bool out) {
                const int n = last - t;
                const double chi = opacity(c, s, out);
                if (t >= 1) {
                    const double dt = 0.5 * (chi + chi_p) * dz;   // dtau[n]
                    if (n == last - 1) {
                        dtL = dt;
                        const double Bd_min_Ad = 1.0 + 2.0 * (1.0 / dt);
                        G_p = 0.5 * Bd_min_Ad * dt * dt;              // G[last]
                    } else {
                        const double iA = 0.5 * (dt + dt_p) * dt;     // inverse_A[n+1]
                        const double iC = 0.5 * (dt + dt_p) * dt_p;   // inverse_C[n+1]
                        G_p = (1.0 + Prec<double>::rcp(iC) * G_p * i1G_p) * iA;  // G[n+1]
                    }
                    i1G_p = Prec<double>::rcp(1.0 + G_p);
                    const int w = n + 1 - w0;
                    if (w >= 0 && w < W) wref(BW_G, w) = G_p;
                    dt_p = dt;
                }
                chi_p = chi;
            });
        }
        // per slot of the pair: this lane's element, then the wave's sum per transition
        int kmin = k, kmax = k < L ? k : -1;
        for (int off = WAVE / 2; off >= 1; off >>= 1) {
            kmin = min(kmin, __shfl_xor(kmin, off));
            kmax = max(kmax, __shfl_xor(kmax, off));
        }
        const double factor = A_k * (w_z * 2.0 * a.ray_w[r]);
        for (int i = 0; i < 2 * nb; i++) {
            const bool upper = i >= nb;
            const int m = upper ? i - nb + 1 : i + 1;
            double v = 0.0;
            if (!upper && nB >= 1 && nA >= m) {
                // L_lower[m][n] = T^-1[n_ar+1][n], n = n_ar - m
                const int n = nA - m, w = n - w0;
                double T;
                if (n == 0) {
                    const double B0_min_C0 = 1.0 + 2.0 * (1.0 / dt0);
                    const double B0 = B0_min_C0 + 2.0 * (1.0 / dt0) * (1.0 / dt0);
                    const double G1 = wref(BW_G, 1 - w0);
                    T = (1.0 + G1) / (B0_min_C0 + B0 * G1);
                } else {
                    const double G1 = wref(BW_G, w + 1);
                    T = wref(BW_IC, w) / (wref(BW_F, w) + G1 * (1.0 / (1.0 + G1)));
                }
                for (int q = n + 1; q <= nA + 1; q++) T = T * (1.0 / (1.0 + wref(BW_G, q - w0)));
                v = factor * (wref(BW_FP, w) * T / wref(BW_CHI, w));
            } else if (upper && nB > 2 * m) {
                // L_upper[m][n] = T^-1[n][n_ar+2m+1], n = n_ar + m
                const int n = nA + m, j = nA + 2 * m + 1;
                double T;
                if (j == last) {
                    const double Fl = wref(BW_F, j - 1 - w0);
                    const double Bd_min_Ad = 1.0 + 2.0 * (1.0 / dtL);
                    const double Bd = Bd_min_Ad + 2.0 * (1.0 / dtL) * (1.0 / dtL);
                    T = (1.0 + Fl) / (Bd_min_Ad + Bd * Fl);
                } else {
                    const double Gj = wref(BW_G, j + 1 - w0);
                    T = wref(BW_IC, j - w0) / (wref(BW_F, j - w0) + Gj * (1.0 / (1.0 + Gj)));
                }
                for (int q = j - 1; q >= n; q--) T = T * (1.0 / (1.0 + wref(BW_F, q - w0)));
                v = factor * (wref(BW_FP, n - w0) * T / wref(BW_CHI, n - w0));
            }
            if (k >= L) v = 0.0;
            double mine = 0.0;
            for (int kk = kmin; kk <= kmax; kk++) {
                double x = (k == kk) ? v : 0.0;
                for (int off = WAVE / 2; off >= 1; off >>= 1) x += __shfl_xor(x, off);
                x = __shfl(x, 0);   // lane 0's order for every lane
                if (lane == kk) mine = x;
            }
            if (lane < L) part[((int64_t)r * 2 * nb + i) * L + lane] = mine;
        }
    }
}

// Band sums per (origin, slot, transition): the waves of an origin in chunk order.
__global__ void k_band_sum(int64_t N, int64_t SL, int nchunk, const double *__restrict__ part,
                           double *__restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * SL) return;
    const int64_t o = t / SL, i = t - o * SL;
    double acc = 0.0;
    for (int c = 0; c < nchunk; c++) acc += part[(o * nchunk + c) * SL + i];
    out[t] = acc;
}

// Largest chunk of staged steps that fits the LDS budget with the cells'
// line rows, at most CH_MAX.
template <typename real, bool X64 = false>
int max_chunk(const mgr_ctx *ctx)
{
    const size_t per_entry = sizeof(step_t<real, X64>) + sizeof(typename Prec<real>::row_t) * ctx->L + 64;
    const int ch = (int)(LDS_LIMIT / per_entry) - 1;
    return ch < CH_MAX ? ch : CH_MAX;
}

// Arguments shared by the radiation-field and image kernels: model tables,
// rows, candidate-line bounds for shifts in [shift_min, shift_max], launch
// shape (nwg waves of 64 bins per workgroup, chunks of ch staged steps) over
// origins [o_begin, o_begin + o_count).  The trace pointers are the caller's.
template <typename real, bool X64 = false>
int make_args(mgr_ctx *ctx, const typename Prec<real>::row_t *rows, int nwg, int ch, int64_t o_begin,
              int64_t o_count, double shift_min, double shift_max, RFArgs<real> &a)
{
    a = RFArgs<real>();
    a.N = ctx->N;
    a.cell_model = ctx->d_cell_model;
    a.cell_qbase = ctx->d_cell_qbase;
    a.m_R = ctx->d_m_R;
    a.m_rbase = ctx->d_m_rbase;
    a.m_active = ctx->d_m_active;
    a.F = (int)ctx->F;
    a.L = (int)ctx->L;
    a.L_magic = 0xFFFFFFFFull / (uint64_t)ctx->L + 1;
    a.nchunk = (int)((ctx->F + WAVE - 1) / WAVE);
    a.nwg = nwg;
    a.ngroups = (a.nchunk + a.nwg - 1) / a.nwg;
    a.o_begin = (int32_t)o_begin;
    a.pair_stride = 1;
    a.tb = ctx->tables;
    a.nu = ctx->nu;
    a.rows = rows;
    a.cell2bdy = ctx->cell2bdy;
    a.antipod = ctx->d_antipod;
    a.I_bdy = ctx->I_bdy;
    a.ray_w = ctx->ray_w;
    a.dshift_max = ctx->dshift_max;
    // A line can be inside the window of bin nu_f at some point only if
    // nu_l (1 - 11 pw_max) <= s_max nu_f and nu_l (1 + 11 pw_max) >= s_min nu_f;
    // the margins cover rounding of the shifts and widths.
    const double w = LINE_WINDOW * ctx->pw_max;
    a.cand_lo = (1.0 + w) * (1.0 + 1.0e-9) / shift_min;   // line * cand_lo < nu_f: below
    a.cand_hi = (1.0 - w) * (1.0 - 1.0e-9) / shift_max;   // line * cand_hi <= nu_f: candidate
    a.lstart_threshold = a.cand_lo;
    a.J = ctx->J;
    a.Lam = ctx->Lam;
    a.f_k = ctx->f_k;
    a.f_z = ctx->f_z;
    a.qw = ctx->qweights;
    a.A_k = ctx->A;
    a.freq_k = ctx->frequency;
    a.pw = ctx->pw;
    a.Lpart = ctx->Lpart;
    a.err_flag = ctx->err_flag;
    if (nwg < 1 || nwg > MAX_WAVES || o_begin < 0 || o_count < 1 || o_begin + o_count > ctx->N)
        return ctx->fail(MGR_ERR_ARGUMENT, "radiation-field launch shape out of range");
    const int64_t groups = o_count * a.ngroups;
    if (groups > INT32_MAX - 8) return ctx->fail(MGR_ERR_UNSUPPORTED, "too many workgroups");
    a.grid_real = (int)groups;
    if (ch < 2 || ch > max_chunk<real, X64>(ctx)) return ctx->fail(MGR_ERR_UNSUPPORTED, "line rows do not fit the LDS budget");
    a.ch = ch;
    a.row_offset = ((sizeof(step_t<real, X64>) * (ch + 1) + 63) / 64) * 64;
    a.line_offset = a.row_offset + sizeof(typename Prec<real>::row_t) * (ch + 1) * ctx->L;
    return MGR_OK;
}

template <typename real>
size_t lds_bytes(const mgr_ctx *ctx, const RFArgs<real> &a)
{
    return a.line_offset + (sizeof(real) == 8 ? sizeof(double) * ctx->L : 0);
}

template <typename real>
unsigned grid_of(const RFArgs<real> &a)
{
    return (unsigned)(((int64_t)a.grid_real + 7) / 8) * 8;
}

// Launch over origins [o_begin, o_begin + o_count) with nwg waves (64 bins
// each) per workgroup and chunks of ch staged steps.
template <typename real>
int launch_rf(mgr_ctx *ctx, const typename Prec<real>::row_t *rows, int nwg, int ch, int64_t o_begin,
              int64_t o_count)
{
    RFArgs<real> a;
    int rc = make_args<real>(ctx, rows, nwg, ch, o_begin, o_count, ctx->shift_min, ctx->shift_max, a);
    if (rc) return rc;
    a.seg_off = ctx->seg_off;
    a.nseg = ctx->nseg;
    a.npts = ctx->npts;
    a.seg_cell = ctx->seg_cell;
    a.seg_shift = ctx->seg_shift;
    a.seg_dZ = ctx->seg_dZ;
    a.seg_f32 = ctx->seg_f32;
    a.seg_hdz32 = ctx->seg_hdz32;
    a.far_cell = ctx->far_cell;
    a.far_shift = ctx->far_shift;
    a.origin_shift = nullptr;
    a.pair_stride = ctx->rf_pair_stride;
    a.u_out = ctx->u_store;
    // the out-of-window bits of the kernel's precision (each exact for its own profile)
    if constexpr (sizeof(real) == 4)
        a.step_mask = (ctx->masks_valid && ctx->rf_window_skip) ? ctx->step_mask : nullptr;
    else
        a.step_mask = (ctx->masks64_valid && ctx->rf_window_skip) ? ctx->step_mask64 : nullptr;
    if (sizeof(real) == 4 && a.step_mask && ctx->hdz32_max > 0.0f) {
        // An out-of-window sub-point closes its row with c = chi_floor dz'_a + chi_floor dz'_b (each
        // product rounded), at most cb = 2 chi_floor dz'_max (1 + 2^-20).  The no-op needs
        // cb < y 2^-25 (then c is below half an ulp of y) and cb y < 2^-25 (then 1 + c y rounds to 1).
        const double cb = 2.0 * (double)(float)(CHI_FLOOR * CHI_SCALE) * (double)ctx->hdz32_max * (1.0 + 0x1p-20);
        const double lo = cb * 0x1p25 * (1.0 + 0x1p-20), hi = 0x1p-25 / cb * (1.0 - 0x1p-20);
        if (lo < hi && lo < 3.0e38 && hi > 1.0e-37) {
            a.noop_lo = std::nextafter((float)lo, INFINITY);   // float(lo) may round below lo
            a.noop_hi = std::nextafter((float)hi, 0.0f);
        }
    }
    a.root_major = ctx->root_major;
    a.far32 = (sizeof(real) == 4 && ctx->far_cache && ctx->far_valid[0]) ? ctx->far_I32 : nullptr;
    a.far64 = (sizeof(real) == 8 && ctx->far_cache && ctx->far_valid[1]) ? ctx->far_I64 : nullptr;
    // FP32: one candidate line per bin when the lines' candidate ranges are disjoint
    bool one = false;
    if constexpr (sizeof(real) == 4) {
        one = true;
        for (int l = 0; l + 1 < ctx->h_tables.L; l++)
            if (!(ctx->h_tables.line[l] * a.cand_lo <= ctx->h_tables.line[l + 1] * a.cand_hi)) one = false;
        if (one) {
            hipLaunchKernelGGL((k_radiation_field<real, true>), dim3(grid_of(a)), dim3(WAVE * a.nwg),
                               lds_bytes(ctx, a), ctx->stream, a);
            ctx->last_rf_kernel = "k_radiation_field<float, true>";
        }
    }
    if (!one) {
        hipLaunchKernelGGL((k_radiation_field<real, false>), dim3(grid_of(a)), dim3(WAVE * a.nwg),
                           lds_bytes(ctx, a), ctx->stream, a);
        ctx->last_rf_kernel = sizeof(real) == 4 ? "k_radiation_field<float, false>" : "k_radiation_field<double, false>";
    }
    MGR_LAUNCH_CHECK();
    return MGR_OK;
}

// Image kernel over every origin, on the rest-frame trace of ctx->img.
template <typename real, bool X64>
int launch_image(mgr_ctx *ctx, const typename Prec<real>::row_t *rows)
{
    const int nchunk = (int)((ctx->F + WAVE - 1) / WAVE);
    const mgr_ctx::ImageTrace &g = ctx->img;
    RFArgs<real> a;
    int rc = make_args<real, X64>(ctx, rows, std::min(nchunk, MAX_WAVES), max_chunk<real, X64>(ctx), 0,
                                  ctx->N, g.shift_min, g.shift_max, a);
    if (rc) return rc;
    a.step_mask = nullptr;   // the masks index the co-moving cache, not the image's trace
    a.seg_off = g.seg_off;
    a.nseg = g.nseg;
    a.npts = g.npts;
    a.seg_cell = g.seg_cell;
    a.seg_shift = g.seg_shift;
    a.seg_dZ = g.seg_dZ;
    a.seg_f32 = g.seg_f32;
    a.seg_hdz32 = g.seg_hdz32;
    a.far_cell = g.far_cell;
    a.far_shift = g.far_shift;
    a.origin_shift = g.origin_shift;
    hipLaunchKernelGGL((k_image<real, X64>), dim3(grid_of(a)), dim3(WAVE * a.nwg), lds_bytes(ctx, a), ctx->stream, a,
                       g.I_m, g.I_p);
    MGR_LAUNCH_CHECK();
    return MGR_OK;
}

}  // namespace

namespace mgr {

// Band elements of Lambda* for the current populations (n_off_diag >= 2):
// k_lambda_band into per-wave shares, then their sums per slot and transition.
int launch_band(mgr_ctx *ctx)
{
    if (ctx->nband <= 0) return MGR_OK;
    RFArgs<double> a;
    int rc = make_args<double>(ctx, ctx->row64, 1, 2, 0, ctx->N, ctx->shift_min, ctx->shift_max, a);
    if (rc) return rc;
    a.seg_off = ctx->seg_off;
    a.nseg = ctx->nseg;
    a.npts = ctx->npts;
    a.seg_cell = ctx->seg_cell;
    a.seg_shift = ctx->seg_shift;
    a.seg_dZ = ctx->seg_dZ;
    a.nband = ctx->nband;
    a.band_S = ctx->band_S;
    a.band_part = ctx->band_part;
    a.step_mask = (ctx->masks64_valid && ctx->rf_window_skip) ? ctx->step_mask64 : nullptr;
    const size_t lds = ((size_t)BW_ARRAYS * 3 * (ctx->nband + 1) * WAVE + ctx->L) * sizeof(double);
    hipLaunchKernelGGL(k_lambda_band, dim3(grid_of(a)), dim3(WAVE), lds, ctx->stream, a);
    MGR_LAUNCH_CHECK();
    const int64_t SL = ctx->band_S * ctx->L;
    hipLaunchKernelGGL(k_band_sum, dim3((unsigned)((ctx->N * SL + 255) / 256)), dim3(256), 0, ctx->stream, ctx->N,
                       SL, a.nchunk, ctx->band_part, ctx->band_Ls);
    MGR_LAUNCH_CHECK();
    return MGR_OK;
}

template <typename real>
static int launch_rf_shape(mgr_ctx *ctx, int nwg, int ch, int64_t o_begin, int64_t o_count)
{
    if constexpr (sizeof(real) == 8)
        return launch_rf<double>(ctx, ctx->row64, nwg, ch, o_begin, o_count);
    else
        return launch_rf<float>(ctx, ctx->row32, nwg, ch, o_begin, o_count);
}

struct TuneCand { int nwg, ch; float ms; };

// Launch-shape autotuner: waves (64-bin chunks) per workgroup and steps per
// staged chunk, timed on up to 16384 origins of the model; the fastest shape
// (or the one with the fewest waves among near-ties) is kept for the context
// and precision.  Runs
// outside graph capture, before the first radiation field of a precision.
template <typename real>
static int tune(mgr_ctx *ctx, int pi)
{
    const int nchunk = (int)((ctx->F + WAVE - 1) / WAVE);
    const int chmax = max_chunk<real>(ctx);
    if (chmax < 2) return ctx->fail(MGR_ERR_UNSUPPORTED, "line rows do not fit the LDS budget");
    std::vector<int> nwgs;
    for (int w = 1; w <= MAX_WAVES; w++)
        if (w == 1 || w == nchunk || (nchunk % w) == 0 || w == std::min(nchunk, MAX_WAVES)) nwgs.push_back(w);
    std::vector<int> chs;
    for (int c : {16, 32, 64})
        if (c <= chmax) chs.push_back(c);
    if (chs.empty() || chs.back() != chmax) chs.push_back(chmax);
    // up to 16384 origins in four slices spread over the model (at 1/8, 3/8,
    // 5/8 and 7/8 of its cells), so that the timing sees inner and outer cells
    const int64_t cnt = std::min<int64_t>(ctx->N, 16384);
    const int nslice = cnt >= 4096 ? 4 : 1;
    const int64_t per = cnt / nslice;
    int64_t os[4];
    for (int k = 0; k < nslice; k++)
        os[k] = nslice == 1 ? (ctx->N - cnt) / 2
                            : std::min<int64_t>(std::max<int64_t>((2 * k + 1) * ctx->N / 8 - per / 2, 0), ctx->N - per);
    auto launch_slices = [&](int w, int c) {
        int r = MGR_OK;
        for (int k = 0; k < nslice && !r; k++) r = launch_rf_shape<real>(ctx, w, c, os[k], per);
        return r;
    };
    int best_w = std::min(nchunk, MAX_WAVES), best_c = chmax;
    float best = 1e30f;
    hipEvent_t e0, e1;
    MGR_HIP(hipEventCreate(&e0));
    MGR_HIP(hipEventCreate(&e1));
    int rc = launch_slices(best_w, best_c);   // warm-up
    ctx->tune_log.clear();
    for (int w : nwgs) {
        for (int c : chs) {
            if (rc) break;
            float ms = 0.0f;
            if (hipEventRecord(e0, ctx->stream) != hipSuccess) { rc = MGR_ERR_DEVICE; break; }
            rc = launch_slices(w, c);
            if (rc) break;
            if (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                rc = MGR_ERR_DEVICE;
                break;
            }
            ctx->tune_log.push_back({w, c, ms});
            if (ms < best) { best = ms; best_w = w; best_c = c; }
        }
    }
    if (rc) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return rc == MGR_ERR_DEVICE ? ctx->fail(MGR_ERR_DEVICE, "radiation-field tuning failed") : rc;
    }
    // Shapes within 2 % of the fastest are a tie on the slice; the one with
    // fewer waves per workgroup schedules the end of a full launch more evenly
    // (measured ~3 % faster on full launches where the slice could not tell).
    float pick = best;
    for (const auto &t : ctx->tune_log)
        if (t.ms <= best * 1.02f && (t.nwg < best_w || (t.nwg == best_w && t.ms < pick))) {
            best_w = t.nwg;
            best_c = t.ch;
            pick = t.ms;
        }
    // The slices do not always rank shapes as full launches do: when a full
    // launch is cheap (under ~0.4 s by the slice timing), the fastest slice
    // shape of each wave count within 5 % of the best is timed on the whole
    // model once, and the fastest full launch is kept.
    const double full_est = (double)best * (double)ctx->N / (double)(per * nslice);
    if (nslice > 1 && full_est < 400.0) {
        std::vector<TuneCand> cands;
        for (int w : nwgs) {
            float bw = 1e30f;
            int bc = 0;
            for (const auto &t : ctx->tune_log)
                if (t.nwg == w && t.ms < bw) { bw = t.ms; bc = t.ch; }
            if (bc && bw <= best * 1.05f) cands.push_back({w, bc, 0.0f});
        }
        if (cands.size() > 1) {
            float full_best = 1e30f;
            for (auto &cd : cands) {
                float ms = 0.0f;
                if (hipEventRecord(e0, ctx->stream) != hipSuccess) { rc = MGR_ERR_DEVICE; break; }
                rc = launch_rf_shape<real>(ctx, cd.nwg, cd.ch, 0, ctx->N);
                if (rc) break;
                if (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                    rc = MGR_ERR_DEVICE;
                    break;
                }
                // ties within 1 % go to the fewer waves (the candidates come in that order)
                if (ms < full_best * 0.99f) {
                    full_best = ms;
                    best_w = cd.nwg;
                    best_c = cd.ch;
                }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc == MGR_ERR_DEVICE ? ctx->fail(MGR_ERR_DEVICE, "radiation-field tuning failed") : rc;
    ctx->rf_nwg[pi] = best_w;
    ctx->rf_ch[pi] = best_c;
    return MGR_OK;
}

int tune_radiation_field(mgr_ctx *ctx, int precision)
{
    const int pi = precision == MGR_FP64 ? 1 : 0;
    if (ctx->rf_nwg[pi] > 0) return MGR_OK;
    if (ctx->rf_autotune) return pi ? tune<double>(ctx, 1) : tune<float>(ctx, 0);
    const int nchunk = (int)((ctx->F + WAVE - 1) / WAVE);
    ctx->rf_nwg[pi] = std::min(nchunk, MAX_WAVES);
    ctx->rf_ch[pi] = pi ? max_chunk<double>(ctx) : max_chunk<float>(ctx);
    return MGR_OK;
}

int launch_rf_kernel(mgr_ctx *ctx, int precision)
{
    const int pi = precision == MGR_FP64 ? 1 : 0;
    if (ctx->rf_nwg[pi] <= 0) return ctx->fail(MGR_ERR_STATE, "radiation-field launch shape not chosen");
    return pi ? launch_rf_shape<double>(ctx, ctx->rf_nwg[1], ctx->rf_ch[1], 0, ctx->N)
              : launch_rf_shape<float>(ctx, ctx->rf_nwg[0], ctx->rf_ch[0], 0, ctx->N);
}

int launch_radiation_field(mgr_ctx *ctx, int precision)
{
    int rc = launch_rf_kernel(ctx, precision);
    if (!rc) rc = launch_band(ctx);
    if (rc) return rc;
    return launch_lambda(ctx);   // update_Lambda's sums over the line bins
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
    int rc0 = set_active(ctx, true);
    if (!rc0) rc0 = launch_step_masks(ctx, precision == MGR_FP64);
    if (!rc0 && ctx->nband > 0) rc0 = launch_step_masks(ctx, true);   // the band kernel's
    if (!rc0) rc0 = launch_far_cache(ctx, precision);
    if (!rc0) rc0 = tune_radiation_field(ctx, precision);
    if (rc0) return rc0;
    hipEvent_t e0, e1;
    MGR_HIP(hipEventCreate(&e0));
    MGR_HIP(hipEventCreate(&e1));
    MGR_HIP(hipMemsetAsync(ctx->err_flag, 0, sizeof(int32_t), ctx->stream));
    // the events bracket the radiation-field kernel alone
    MGR_HIP(hipEventRecord(e0, ctx->stream));
    int rc = launch_rf_kernel(ctx, precision);
    if (!rc) {
        MGR_HIP(hipEventRecord(e1, ctx->stream));
        rc = launch_band(ctx);
        if (!rc) rc = launch_lambda(ctx);
    }
    if (rc) { (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); return rc; }
    MGR_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    ctx->last_rf_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    MGR_HIP(hipStreamSynchronize(ctx->stream));   // the band elements and Lambda* sums too
    int32_t errv = 0;
    MGR_HIP(hipMemcpy(&errv, ctx->err_flag, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (errv) return ctx->fail(MGR_ERR_TRACE, "a ray pair ended off the boundary or an interior origin has an empty ray");
    ctx->J_set = true;
    return MGR_OK;
}

// The warm start's first pass: a radiation field over every pair_stride-th
// ray pair of each origin (weights scaled by pair_stride) and J_eff.
extern "C" int mgr_first_pass(mgr_ctx *ctx, int precision, int pair_stride)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    if (precision != MGR_FP32 && precision != MGR_FP64) return ctx->fail(MGR_ERR_ARGUMENT, "precision must be MGR_FP32 or MGR_FP64");
    if (pair_stride < 1) return ctx->fail(MGR_ERR_ARGUMENT, "pair_stride must be at least 1");
    if (!ctx->spectral_set) return ctx->fail(MGR_ERR_STATE, "compute_spectral_discretisation first");
    if (!ctx->bdy_set) return ctx->fail(MGR_ERR_STATE, "compute_boundary_intensities first");
    if (!ctx->pops_set) return ctx->fail(MGR_ERR_STATE, "level populations not set");
    MGR_HIP(hipSetDevice(ctx->device));
    if (!ctx->traced) {
        int rc = launch_trace(ctx);
        if (rc) return rc;
    }
    int rc = set_active(ctx, true);
    if (!rc) rc = launch_step_masks(ctx, precision == MGR_FP64);
    if (!rc && ctx->nband > 0) rc = launch_step_masks(ctx, true);   // the band kernel's
    if (!rc) rc = launch_far_cache(ctx, precision);
    if (!rc) rc = tune_radiation_field(ctx, precision);
    if (rc) return rc;
    MGR_HIP(hipMemsetAsync(ctx->err_flag, 0, sizeof(int32_t), ctx->stream));
    ctx->rf_pair_stride = pair_stride;
    rc = launch_rf_kernel(ctx, precision);
    ctx->rf_pair_stride = 1;
    if (!rc) rc = launch_band(ctx);
    if (!rc) rc = launch_lambda(ctx);
    if (!rc) rc = launch_jeff(ctx);
    if (rc) return rc;
    MGR_HIP(hipStreamSynchronize(ctx->stream));
    int32_t errv = 0;
    MGR_HIP(hipMemcpy(&errv, ctx->err_flag, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (errv) return ctx->fail(MGR_ERR_TRACE, "a ray pair ended off the boundary or an interior origin has an empty ray");
    ctx->J_set = true;
    return MGR_OK;
}

extern "C" int mgr_radiation_field_shape(mgr_ctx *ctx, int precision, int *waves_per_group, int *steps_per_chunk,
                                         int capacity, int *n_tried, int *tried_waves, int *tried_steps,
                                         float *tried_ms)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    if (precision != MGR_FP32 && precision != MGR_FP64) return ctx->fail(MGR_ERR_ARGUMENT, "bad precision");
    const int pi = precision == MGR_FP64 ? 1 : 0;
    if (waves_per_group) *waves_per_group = ctx->rf_nwg[pi];
    if (steps_per_chunk) *steps_per_chunk = ctx->rf_ch[pi];
    const int n = (int)ctx->tune_log.size();
    if (n_tried) *n_tried = n;
    for (int i = 0; i < n && i < capacity; i++) {
        if (tried_waves) tried_waves[i] = ctx->tune_log[i].nwg;
        if (tried_steps) tried_steps[i] = ctx->tune_log[i].ch;
        if (tried_ms) tried_ms[i] = ctx->tune_log[i].ms;
    }
    return MGR_OK;
}

extern "C" int mgr_set_radiation_field_shape(mgr_ctx *ctx, int precision, int waves_per_group, int steps_per_chunk)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    if (precision != MGR_FP32 && precision != MGR_FP64) return ctx->fail(MGR_ERR_ARGUMENT, "bad precision");
    const int pi = precision == MGR_FP64 ? 1 : 0;
    if (waves_per_group == 0 && steps_per_chunk == 0) {
        ctx->rf_nwg[pi] = ctx->rf_ch[pi] = 0;   // the next radiation field tunes again
    } else {
        const int chmax = pi ? max_chunk<double>(ctx) : max_chunk<float>(ctx);
        if (waves_per_group < 1 || waves_per_group > MAX_WAVES)
            return ctx->fail(MGR_ERR_ARGUMENT, "waves_per_group must be 1.." + std::to_string(MAX_WAVES));
        if (steps_per_chunk < 2 || steps_per_chunk > chmax)
            return ctx->fail(MGR_ERR_ARGUMENT, "steps_per_chunk must be 2.." + std::to_string(chmax) +
                                                   " for this model's line rows");
        ctx->rf_nwg[pi] = waves_per_group;
        ctx->rf_ch[pi] = steps_per_chunk;
    }
    // the ALI graphs were captured with the previous shape
    if (ctx->graph_precision == precision) {
        MGR_HIP(hipSetDevice(ctx->device));
        if (ctx->graph_normal) (void)hipGraphExecDestroy(ctx->graph_normal);
        if (ctx->graph_ng) (void)hipGraphExecDestroy(ctx->graph_ng);
        ctx->graph_normal = ctx->graph_ng = nullptr;
        ctx->graph_precision = -1;
    }
    return MGR_OK;
}

extern "C" int mgr_compute_image(mgr_ctx *ctx, int64_t ray, int precision, double *I_m, double *I_p)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    if (precision != MGR_FP32 && precision != MGR_FP64) return ctx->fail(MGR_ERR_ARGUMENT, "precision must be MGR_FP32 or MGR_FP64");
    // sim_radiation.tpp:353-358
    if (!ctx->spectral_set || !ctx->spectral_image)
        return ctx->fail(MGR_ERR_STATE, "Spectral discretisation was not set for Images");
    if (!ctx->bdy_set) return ctx->fail(MGR_ERR_STATE, "compute_boundary_intensities first");
    if (!ctx->pops_set) return ctx->fail(MGR_ERR_STATE, "level populations not set");
    if (ctx->M != 1) return ctx->fail(MGR_ERR_UNSUPPORTED, "images are per model (batch of one)");
    if (ray < 0 || ray >= ctx->R / 2) return ctx->fail(MGR_ERR_ARGUMENT, "image ray must be in [0, nrays/2)");
    MGR_HIP(hipSetDevice(ctx->device));
    int rc = launch_trace_image(ctx, ray);
    if (rc) return rc;
    mgr_ctx::ImageTrace &g = ctx->img;
    const int64_t NF = ctx->N * ctx->F;
    if (!g.I_m) {
        MGR_HIP(hipMalloc((void **)&g.I_m, NF * sizeof(double)));
        MGR_HIP(hipMalloc((void **)&g.I_p, NF * sizeof(double)));
    }
    MGR_HIP(hipMemsetAsync(ctx->err_flag, 0, sizeof(int32_t), ctx->stream));
    // FP32: the sweep and profiles in FP32, shifts and profile arguments in FP64
    rc = precision == MGR_FP64 ? launch_image<double, false>(ctx, ctx->row64)
                               : launch_image<float, true>(ctx, ctx->row32);
    if (rc) return rc;
    MGR_HIP(hipStreamSynchronize(ctx->stream));
    int32_t errv = 0;
    MGR_HIP(hipMemcpy(&errv, ctx->err_flag, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (errv) return ctx->fail(MGR_ERR_TRACE, "an image ray ended off the boundary or an interior origin has an empty ray");
    if (I_m) MGR_HIP(hipMemcpy(I_m, g.I_m, NF * sizeof(double), hipMemcpyDeviceToHost));
    if (I_p) MGR_HIP(hipMemcpy(I_p, g.I_p, NF * sizeof(double), hipMemcpyDeviceToHost));
    return MGR_OK;
}

// Radiation::u (radiation.cpp:117-126): with scattering every radiation
// field keeps each ray pair's u at the origin (sim_radiation.tpp:226-237,
// 258-262).  Allocates [R/2][N*F] doubles on the device; the ALI graphs are
// rebuilt so that they write it too.
extern "C" int mgr_set_scattering(mgr_ctx *ctx, int use_scattering)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    MGR_HIP(hipSetDevice(ctx->device));
    MGR_HIP(hipStreamSynchronize(ctx->stream));
    if (use_scattering && !ctx->u_store) {
        const size_t n = (size_t)(ctx->R / 2) * ctx->N * ctx->F;
        MGR_HIP(hipMalloc((void **)&ctx->u_store, std::max<size_t>(n, 1) * sizeof(double)));
        MGR_HIP(hipMemset(ctx->u_store, 0, std::max<size_t>(n, 1) * sizeof(double)));
    } else if (!use_scattering && ctx->u_store) {
        (void)hipFree(ctx->u_store);
        ctx->u_store = nullptr;
    }
    ctx->graph_precision = -1;   // captured launches hold the old pointer
    return MGR_OK;
}

namespace mgr {

// Builds the far-end intensity cache of a precision when it is missing or
// stale (once per trace, frequency grid and boundary intensities).  It is
// skipped, and the kernel computes the intensities itself, when the cache
// would take more than three quarters of the free device memory or an entry
// has no boundary cell at its far end (the kernel then reports that error).
int launch_far_cache(mgr_ctx *ctx, int precision)
{
    const int pi = precision == MGR_FP64 ? 1 : 0;
    if (!ctx->far_cache || ctx->far_valid[pi] || ctx->far_bad[pi] || !ctx->traced || !ctx->spectral_set ||
        !ctx->bdy_set)
        return MGR_OK;
    void **buf = pi ? (void **)&ctx->far_I64 : (void **)&ctx->far_I32;
    const size_t bytes = (size_t)ctx->NQ * (size_t)ctx->F * (pi ? sizeof(double) : sizeof(float));
    if (!*buf) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess || bytes > fr / 4 * 3 ||
            hipMalloc(buf, std::max<size_t>(bytes, 8)) != hipSuccess) {
            (void)hipGetLastError();
            *buf = nullptr;
            ctx->far_bad[pi] = true;
            return MGR_OK;
        }
    }
    ctx->graph_precision = -1;   // captured launches may have read no cache, or a stale one
    RFArgs<float> a = RFArgs<float>();
    a.N = ctx->N;
    a.F = (int)ctx->F;
    a.cell_model = ctx->d_cell_model;
    a.cell_qbase = ctx->d_cell_qbase;
    a.m_R = ctx->d_m_R;
    a.nu = ctx->nu;
    a.cell2bdy = ctx->cell2bdy;
    a.I_bdy = ctx->I_bdy;
    a.far_cell = ctx->far_cell;
    a.far_shift = ctx->far_shift;
    const int64_t entries = ctx->N * ctx->R;   // one wave each, four to a workgroup
    const dim3 grid((unsigned)((entries + 3) / 4));
    MGR_HIP(hipMemsetAsync(ctx->err_flag, 0, sizeof(int32_t), ctx->stream));
    if (pi)
        hipLaunchKernelGGL(k_far_cache<double>, grid, dim3(256), 0, ctx->stream, a, (int)ctx->R, ctx->far_I64,
                           ctx->err_flag);
    else
        hipLaunchKernelGGL(k_far_cache<float>, grid, dim3(256), 0, ctx->stream, a, (int)ctx->R, ctx->far_I32,
                           ctx->err_flag);
    MGR_LAUNCH_CHECK();
    int32_t errv = 0;
    MGR_HIP(hipStreamSynchronize(ctx->stream));
    MGR_HIP(hipMemcpy(&errv, ctx->err_flag, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (errv) ctx->far_bad[pi] = true;
    else ctx->far_valid[pi] = true;
    return MGR_OK;
}

}  // namespace mgr

extern "C" int mgr_set_far_cache(mgr_ctx *ctx, int enable)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    MGR_HIP(hipSetDevice(ctx->device));
    MGR_HIP(hipStreamSynchronize(ctx->stream));
    ctx->far_cache = enable != 0;
    if (!ctx->far_cache) {
        if (ctx->far_I32) (void)hipFree(ctx->far_I32);
        if (ctx->far_I64) (void)hipFree(ctx->far_I64);
        ctx->far_I32 = nullptr;
        ctx->far_I64 = nullptr;
    }
    ctx->far_invalidate();
    ctx->graph_precision = -1;   // captured launches hold the old choice
    return MGR_OK;
}

extern "C" int mgr_set_window_skip(mgr_ctx *ctx, int enable)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    ctx->rf_window_skip = enable != 0;
    ctx->graph_precision = -1;   // captured launches hold the old choice
    return MGR_OK;
}

extern "C" int mgr_get_u(mgr_ctx *ctx, double *u)
{
    if (!ctx || !u) return MGR_ERR_ARGUMENT;
    if (!ctx->u_store) return ctx->fail(MGR_ERR_STATE, "scattering is off (mgr_set_scattering)");
    MGR_HIP(hipSetDevice(ctx->device));
    MGR_HIP(hipStreamSynchronize(ctx->stream));
    MGR_HIP(hipMemcpy(u, ctx->u_store, (size_t)(ctx->R / 2) * ctx->N * ctx->F * sizeof(double),
                      hipMemcpyDeviceToHost));
    return MGR_OK;
}
