// mgr_common.hpp -- shared definitions of the gfx950 radiation-field /
// level-population library (device layout, constants, context).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/magritte_hip.h"

namespace mgr {

// Physical constants, src/Tools/constants.hpp:26-63 (the same literal values
// and derivations, so that every derived quantity rounds as the reference's).
constexpr double PI              = 3.14159265358979323846;
constexpr double FOUR_PI         = 4.0 * PI;
constexpr double INVERSE_SQRT_PI = 0.5 * 1.12837916709551257390;   // 0.5*M_2_SQRTPI
constexpr double CC              = 2.99792458E+8;
constexpr double HH              = 6.62607004E-34;
constexpr double KB              = 1.38064852E-23;
constexpr double AMU             = 1.66053904E-27;
constexpr double T_CMB           = 2.7254800;
constexpr double C_SQUARED       = 8.98755179E+16;
constexpr double TWO_KB_OVER_AMU_C_SQUARED = 2.0 * KB / (AMU * C_SQUARED);
constexpr double TWO_HH_OVER_CC_SQUARED    = 2.0 * HH / C_SQUARED;
constexpr double HH_OVER_KB                = HH / KB;
constexpr double HH_OVER_FOUR_PI           = HH / FOUR_PI;

// Line window of get_eta_and_chi (src/Simulation/simulation.tpp:403-404) and
// opacity floor (:397).
constexpr double LINE_WINDOW = 11.0;
constexpr double CHI_FLOOR   = 1.0E-26;

// Exact power-of-two scalings of the radiation-field solve.  Opacities are
// carried as chi * CHI_SCALE and path lengths as dZ / CHI_SCALE (optical depth
// unchanged); intensities and source functions as I * I_SCALE.  In FP32 this
// keeps chi, S, dtau^2 and S*dtau^2 inside the normal range for SI inputs;
// powers of two make the scaling exact in either precision.
constexpr double CHI_SCALE = 73786976294838206464.0;     // 2^66 ~ 7.4e19
constexpr double I_SCALE   = 1125899906842624.0;         // 2^50 ~ 1.1e15

constexpr double SQRT_LOG2E = 1.2011224087864498;   // sqrt(log2(e))
constexpr double SQRT2      = 1.4142135623730951;
// Path-length factor of the solve's scaled interval c = dtau / sqrt(2):
// c = (chi_a + chi_b) * dZ * HDZ_SCALE with chi scaled by CHI_SCALE.
constexpr double HDZ_SCALE  = 0.35355339059327373 / CHI_SCALE;   // 1 / (2 sqrt 2) / CHI_SCALE

constexpr int WAVE = 64;                 // CDNA wavefront width
// Packing of a step's counts (the FP32 trace record and the staged step):
// sub-points n (bits 0-13, at most 10000), the first sub-point of the step's
// own cell `half` (bits 14-26), and one bit per 64-bin chunk
// k < STEP_MASK_CHUNKS (bit 27 + k) set when every bin of the chunk is more
// than MASK_WINDOW line widths from every line at every sub-point of the step:
// the FP32 profile exp2(-x^2 log2 e) is then exactly zero, since x >= 10.5
// gives exp2 of less than -159, below the smallest denormal 2^-149 (zero from
// x > 10.2; the FP32 rounding of x is ~1e-6 widths).
constexpr int STEP_N_MASK = 0x3FFF, STEP_HALF_SHIFT = 14, STEP_HALF_MASK = 0x1FFF, STEP_OUT_SHIFT = 27;
constexpr int STEP_MASK_CHUNKS = 5;
constexpr double MASK_WINDOW = 10.5;
// The FP64 kernel's bits: more than 11 (1 + 1e-6) widths from every line, so
// that get_eta_and_chi's own +-11-width window (LINE_WINDOW) holds no line at
// any sub-point whatever the rounding of the shift (~1e-16): eta = 0 and chi
// at the floor exactly, the values the out-of-window branch uses.
constexpr double MASK_WINDOW64 = 11.0 * (1.0 + 1e-6);
constexpr int MAX_INTERPOLATION = 10000; // geometry.tpp:138

// One neighbour entry of the tracer: the neighbour's index and position, so
// that a step reads its candidates contiguously (one gather less per step).
struct alignas(32) NbRec {
    double  x, y, z;
    int64_t c;
};

// Small per-model line tables, read by every kernel.
constexpr int MAX_SPECIES = 8;    // line-producing species per context
struct LineTables {
    int    L;              // lines of every species (Parameters::nlines)
    int    V;              // levels of every species
    int    Q;              // nquads
    int    F;              // nfreqs = L*Q
    int    S;              // line-producing species
    double inverse_mass;   // species 0
    double line[64];       // sorted line frequencies (Lines::line)
    int    line_index[64]; // line g (species after species) of sorted line l (Lines::line_index)
    // per species: levels, lines, first population (in units of ncells), first
    // line g, inverse mass
    int    sV[MAX_SPECIES], sL[MAX_SPECIES], voff[MAX_SPECIES], loff[MAX_SPECIES];
    double inv_mass[MAX_SPECIES];
    int    line_spec[64];  // species of line g
    // species whose mass sets the width of sorted line l in get_eta_and_chi:
    // Simulation::get_line_width (simulation.tpp:363-373) reads the frequency
    // back-map corresponding_l_for_spec at index line_index[l], a bin index
    int    width_spec[64];
};
constexpr int MAX_LINES = 64;

// Bin held by lane slot i (0 <= i < F = L Q) of the radiation-field kernel:
// the bin itself, or in root-major order root i / L of sorted line i % L (bin
// l Q + z), so that each 64-bin wave holds the same few roots of every line
// and leaves all line windows at the same shifts.
__host__ __device__ inline int lane_bin(int i, int L, int F, bool root_major)
{
    return root_major ? (i % L) * (F / L) + i / L : i;
}

#define MGR_HIP(call)                                                              \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            ctx->fail(MGR_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
            return MGR_ERR_DEVICE;                                                 \
        }                                                                          \
    } while (0)

#define MGR_LAUNCH_CHECK()                                                         \
    do {                                                                           \
        hipError_t e_ = hipGetLastError();                                         \
        if (e_ != hipSuccess) {                                                    \
            ctx->fail(MGR_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e_)); \
            return MGR_ERR_DEVICE;                                                 \
        }                                                                          \
    } while (0)

}  // namespace mgr

// Device-resident model and state (SoA).  Indices follow the reference:
// J(p,f) at f + p*F (radiation.tpp:11-15), populations at i + p*V
// (lineProducingSpecies.tpp:24-30), emissivity/opacity at k + p*L (lines.tpp:19-34).
struct mgr_ctx {
    int         device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int         status_sticky = MGR_OK;

    // sizes (N, NB: totals over the models of a batch; R: the largest ray count)
    int64_t N = 0, R = 0, NB = 0, F = 0, L = 0, V = 0, Q = 0, nspecs = 0;

    // batch of models: model m owns cells [m_cell_off[m], m_cell_off[m+1]) and
    // the (origin, ray) entries [m_q_off[m], m_q_off[m+1]); its rays are the
    // global rays m_rbase[m] .. m_rbase[m] + m_R[m] - 1 (one set per distinct
    // ray list).  A single model is a batch of one.
    int64_t M = 1, NQ = 0;
    std::vector<int64_t> m_cell_off, m_q_off, h_cell_qbase;
    std::vector<int32_t> m_R, m_rbase;
    int32_t *d_cell_model = nullptr;   // [N]
    int64_t *d_cell_qbase = nullptr;   // [N] first (origin, ray) entry of the cell
    int32_t *d_m_R = nullptr, *d_m_rbase = nullptr, *d_m_active = nullptr;   // [M]
    int32_t *h_m_active = nullptr;     // pinned [M]
    // segmented reductions: chunk c covers cells [chunk_cell[c], chunk_cell[c+1])
    // of one model; the chunks of model m are m_chunk_off[m] .. m_chunk_off[m+1]-1
    int64_t nchunks = 0;
    std::vector<int64_t> m_chunk_off;
    int64_t *d_chunk_cell = nullptr;   // [nchunks + 1]
    int64_t *d_m_chunk_off = nullptr;  // [M + 1]
    int64_t *d_m_cell_off = nullptr;   // [M + 1]
    double  *d_mconv = nullptr;        // [M * 4] per-model fraction, mean, max, error word
    double  *h_mconv = nullptr;        // pinned copy
    double  *d_mng = nullptr;          // [M * 4] per-model Ng determinant and coefficients
    std::vector<int64_t> m_iterations; // per model, of the last level-population run
    std::vector<mgr_conv> m_last_conv;
    int64_t ncolpar = 0, ncol_tot = 0, ntmp_tot = 0, ctab_tot = 0;
    // line-producing species: levels, lines, first population block (in
    // units of N), first line, abundance index, first collision partner and
    // partner count of each; populations are held species after species
    int64_t S = 1;
    std::vector<int64_t> sp_V, sp_L, sp_voff, sp_loff, sp_num, sp_cp0, sp_ncp;
    double  pop_prec = 1e-6, max_width_fraction = 0.5;
    int64_t num = 0;

    // host copies
    std::vector<int64_t> antipod;          // [R]
    std::vector<int64_t> boundary2cell;    // [NB]
    std::vector<int64_t> lam_f_off, lam_f; // CSR: sorted f of each transition k
    std::vector<int64_t> h_cp_num;         // [ncolpar] species of each collision partner

    // device: static model
    double  *pos = nullptr, *vel = nullptr;        // [N*3]
    int64_t *nb_off = nullptr;                     // [N+1]
    int32_t *nb = nullptr;                         // [sum n_neighbors]
    mgr::NbRec *nb_rec = nullptr;                  // [sum n_neighbors] neighbour index and position
    int64_t  nb_kmax = 0;                          // longest neighbour list
    double   fwd_tau = 0.0;                        // candidate-mask margin (1e-9 of the extent)
    uint64_t *fwd_mask = nullptr;                  // [N*R] candidate neighbours per (cell, ray), or null
    int64_t  trace_record_steps = 256;             // cells recorded per walk by the counting pass (0: none)
    bool     trace_masks = true;                   // candidate masks on
    uint8_t *is_boundary = nullptr;                // [N]
    int32_t *cell2bdy = nullptr;                   // [N], -1 for interior cells
    int32_t *bdy2cell = nullptr;                   // [NB]
    double  *rays = nullptr, *ray_w = nullptr;     // [R*3], [R]
    int32_t *d_antipod = nullptr;                  // [R]
    double  *temperature = nullptr, *vturb2 = nullptr;   // [N]
    double  *abundance = nullptr;                  // [N*nspecs]
    mgr::LineTables *tables = nullptr;             // device copy
    mgr::LineTables  h_tables;                     // host copy
    double  *A = nullptr, *Bs = nullptr, *Ba = nullptr, *frequency = nullptr; // [L]
    int32_t *irad = nullptr, *jrad = nullptr;      // [L]
    double  *energy = nullptr, *lweight = nullptr; // [V]
    double  *roots = nullptr, *qweights = nullptr; // [S*Q] species after species
    int32_t *cp_num = nullptr, *cp_ncol = nullptr, *cp_ntmp = nullptr;  // [ncolpar]
    int32_t *cp_koff = nullptr, *cp_toff = nullptr, *cp_coff = nullptr; // [ncolpar]
    int32_t *cp_op = nullptr;                      // [ncolpar] 0 n, 1 o, 2 p
    int32_t *icol = nullptr, *jcol = nullptr;      // [ncol_tot]
    double  *tmp = nullptr, *Cd = nullptr, *Ce = nullptr;

    // device: derived per-cell tables
    double  *pw = nullptr;          // [S*N] profile width / line frequency, per species
    double  *pw_cmax = nullptr;     // [N] the largest over species
    double  *dshift_max = nullptr;  // [N]
    double  *W11 = nullptr;         // [N*L] 11 * width of sorted line l
    double  *IW64 = nullptr;        // [N*L] 1 / width
    float   *IW32 = nullptr;        // [N*L]
    // per (cell, sorted line) rows read by the radiation-field kernel:
    // FP32 (1/width, E, K, 0), FP64 (11 width, 1/width, E, K); E and K are the
    // scaled emissivity / opacity prefactors, refreshed with the populations.
    float4  *row32 = nullptr;       // [N*L]
    double4 *row64 = nullptr;       // [N*L]
    double   pw_max = 0.0;

    // device: spectral discretisation
    double  *nu = nullptr;          // [N*F]
    int32_t *nr_line = nullptr;     // [N*L*Q]
    int32_t *f_appears = nullptr, *f_k = nullptr, *f_z = nullptr;   // [F]
    int32_t *d_lam_f_off = nullptr, *d_lam_f = nullptr;             // CSR
    bool     spectral_set = false;
    double  *I_bdy = nullptr;       // [NB*F]
    bool     bdy_set = false;

    // device: populations and line quantities
    double  *pop = nullptr, *prev1 = nullptr, *prev2 = nullptr, *prev3 = nullptr; // [N*V]
    double  *pop_tot = nullptr;     // [S*N]
    double  *emis = nullptr, *opac = nullptr;   // [N*L]
    bool     pops_set = false;

    // device: radiation field
    double  *J = nullptr;           // [N*F]
    double  *Lam = nullptr;         // [N*F]  sum_r 2 w_r Lambda_ar / chi
    double  *Ls = nullptr;          // [N*L]
    // each radiation-field wave's share of the Lambda* sums: [N][ceil(F/64)][L]
    double  *Lpart = nullptr;
    double  *Jlin = nullptr, *Jeff = nullptr, *Jdif = nullptr;  // [N*L]
    // Banded Lambda* (n_off_diag = nband + 1 >= 2, one model): per origin,
    // band_S = (R/2) 2 nband slots, slot (r 2 + side) nband + m - 1 for ray
    // pair r, band m, side 0 the lower band (on the longer ray), 1 the upper
    // band (on the shorter ray).  band_cell: the slot's cell, -1 when the
    // element does not exist; band_Ls: its sum per transition.
    int      nband = 0;
    int64_t  band_S = 0;
    int32_t *band_cell = nullptr;   // [N*band_S], set at trace time
    double  *band_part = nullptr;   // [N][ceil(F/64)][band_S][L] each wave's share
    double  *band_Ls = nullptr;     // [N*band_S*L]
    // slots whose cell is c, in ascending slot order: band_rev[band_rev_off[c] ..]
    int64_t *band_rev_off = nullptr;   // [N+1]
    int64_t *band_rev = nullptr;       // [band_rev_off[N]]
    // the rate equations with band elements: per-cell LU of the block
    // diagonal, the block-Jacobi iterate and the sweep flags
    double  *band_lu = nullptr;     // [N*V*V]
    int32_t *band_piv = nullptr;    // [N*V]
    double  *band_x = nullptr;      // [N*V]
    int32_t *band_flags = nullptr;  // [BAND_SWEEPS + 1]
    unsigned long long *band_change = nullptr;   // [BAND_SWEEPS] largest move of each sweep / n_tot (bits)
    // Radiation::u with scattering: each ray pair's u at the origin,
    // [R/2][N*F] at (r N + p) F + f; null without scattering
    double  *u_store = nullptr;
    int32_t *err_flag = nullptr;    // device error word
    bool     J_set = false;
    float    last_rf_ms = 0.0f;
    const char *last_rf_kernel = "";   // name of the radiation-field kernel last launched
    // radiation-field launch shape per precision (FP32, FP64): waves per
    // workgroup and staged steps per chunk; 0 until chosen
    int      rf_nwg[2] = {0, 0}, rf_ch[2] = {0, 0};
    bool     rf_autotune = true;
    int      rf_pair_stride = 1;    // every k-th ray pair (weights times k): the warm start's first pass only
    int64_t  warm_fallbacks = 0;    // cells that kept their populations in the last warm start
    struct TuneEntry { int nwg, ch; float ms; };
    std::vector<TuneEntry> tune_log;

    // device: trace cache
    int64_t *seg_off = nullptr;     // [N*R + 1]
    int32_t *nseg = nullptr;        // [N*R]
    int32_t *npts = nullptr;        // [N*R]
    int32_t *seg_cell = nullptr;    // [total segments]
    double  *seg_shift = nullptr, *seg_dZ = nullptr;
    // the FP32 kernel's view of every step, split once at trace time:
    // (cell, n | half << STEP_HALF_SHIFT, shift offset s0, shift increment) and dZ * HDZ_SCALE
    int4    *seg_f32 = nullptr;
    float   *seg_hdz32 = nullptr;
    // last emitted point of every (origin, ray): its cell (-1: none) and FP64 shift
    int32_t *far_cell = nullptr;
    double  *far_shift = nullptr;
    int64_t  total_segments = 0;
    uint64_t *shift_bits = nullptr;  // [2] min and max shift over all steps (FP64 bits)
    double   shift_min = 1.0, shift_max = 1.0;
    bool     traced = false;
    // per step, bit k set when no bin of 64-bin chunk k meets a line window at
    // any of the step's sub-points; valid for the current trace and frequencies
    uint8_t *step_mask = nullptr;   // [total segments]
    bool     masks_valid = false;
    uint8_t *step_mask64 = nullptr; // [total segments] the FP64 kernel's bits (MASK_WINDOW64)
    bool     masks64_valid = false;
    float    hdz32_max = 0.0f;      // largest dz' of the cached steps (set with the masks)
    bool     rf_window_skip = true;
    // lane order of the radiation-field kernel (lane_bin), chosen with the
    // masks: root-major when no bin can meet a second line's window
    bool     root_major = false;
    double   root_max = 0.0;            // max |root| of the line quadrature
    // Incoming intensity at the far end of every (origin, ray) entry and bin,
    // (real)(rescaled I_bdy * I_SCALE) at q F + f, in FP32 ([0], float) and FP64
    // ([1], double) for the kernels of each precision; valid for the current
    // trace, frequencies and boundary intensities.  Allocated when it fits three
    // quarters of the free device memory; null otherwise (computed in the kernel).
    float   *far_I32 = nullptr;
    double  *far_I64 = nullptr;
    bool     far_valid[2] = {false, false}, far_bad[2] = {false, false}, far_cache = true;
    void far_invalidate() { far_valid[0] = far_valid[1] = far_bad[0] = far_bad[1] = false; }
    int64_t  sum_ndep = 0, n_pairs_solved = 0;

    // device: rest-frame trace of one image ray pair, entry 2o the antipode
    // side and 2o + 1 the ray side of origin o (same layout as the trace
    // cache), and the image intensities
    struct ImageTrace {
        int64_t  ray = -1;              // traced ray, -1: none
        int64_t *seg_off = nullptr;     // [2N + 1]
        int32_t *nseg = nullptr, *npts = nullptr;   // [2N]
        int32_t *seg_cell = nullptr;
        double  *seg_shift = nullptr, *seg_dZ = nullptr;
        int4    *seg_f32 = nullptr;
        float   *seg_hdz32 = nullptr;
        int32_t *far_cell = nullptr;    // [2N]
        double  *far_shift = nullptr;   // [2N]
        double  *origin_shift = nullptr;   // [2N] rest-frame shift of the origin on each side
        double   shift_min = 1.0, shift_max = 1.0;
        int64_t  total_segments = 0;
        double  *I_m = nullptr, *I_p = nullptr;   // [N*F]
    } img;
    bool     spectral_image = false;    // frequencies set by the image discretisation

    // reductions
    double  *partials = nullptr;    // [nchunks * 8]
    double  *reduced = nullptr;     // [4 (MAX_SPECIES + 1)]: the result, then per species
    double  *h_reduced = nullptr;   // pinned [8]

    // graphs of one ALI iteration
    hipGraphExec_t graph_normal = nullptr, graph_ng = nullptr;
    int graph_precision = -1;

    int fail(int code, const std::string &msg) {
        err = msg;
        return code;
    }
};

namespace mgr {
constexpr int RED_BLOCK = 256;
constexpr int MAX_OFF_DIAG = 8;     // widest Lambda* band (n_off_diag)
constexpr int BAND_SWEEPS = 16;     // block-Jacobi sweeps of the banded rate equations
constexpr int CHUNK_CELLS = 1024;   // cells per reduction chunk

void set_context_free_error(const std::string &msg);

// entry points shared between translation units
int upload_models(mgr_ctx *ctx, const mgr_model_view *models, int64_t nmodels);
int free_model(mgr_ctx *ctx);
int launch_cell_tables(mgr_ctx *ctx);
int launch_emissivity_opacity(mgr_ctx *ctx);
int launch_trace(mgr_ctx *ctx);
int launch_forward_masks(mgr_ctx *ctx);   // candidate neighbours of every (cell, ray), once
int launch_radiation_field(mgr_ctx *ctx, int precision);   // kernel + Lambda*
int launch_far_cache(mgr_ctx *ctx, int precision);   // outside graph capture
int launch_rf_kernel(mgr_ctx *ctx, int precision);
int tune_radiation_field(mgr_ctx *ctx, int precision);   // outside graph capture
int launch_lambda(mgr_ctx *ctx);
int launch_band(mgr_ctx *ctx);         // band elements of Lambda* (nband > 0)
int launch_band_cells(mgr_ctx *ctx);   // band cells and their reverse map (at trace time)
int launch_jeff(mgr_ctx *ctx);
int launch_stateq(mgr_ctx *ctx);
int launch_convergence(mgr_ctx *ctx);
int launch_ng(mgr_ctx *ctx);
int read_convergence(mgr_ctx *ctx, mgr_conv *conv);
int set_active(mgr_ctx *ctx, bool all);   // all: every model of the batch iterates
int launch_trace_image(mgr_ctx *ctx, int64_t ray);   // rest-frame trace of ray and its antipode
int launch_step_masks(mgr_ctx *ctx, bool fp64);   // out-of-window bits of the cached steps (when stale)
void free_image(mgr_ctx *ctx);
}  // namespace mgr
