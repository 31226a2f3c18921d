// mgr_trace.hip -- ray tracer (Geometry::trace_ray<CoMoving>,
// src/Model/Geometry/geometry.tpp:21-95) over the point-cloud neighbour lists.
//
// One thread per (origin, ray).  A counting pass walks every ray, scanning
// only the neighbours that lie ahead along the ray (candidate masks per
// cell and ray, k_forward_masks), and records the cells it visits; the host
// scans the counts, and a filling pass replays the recorded cells (no
// neighbour scans) into one device-resident SoA cache, in blocks of four
// steps, that every later radiation field reuses: geometry and velocities are
// static, so the reference's re-tracing on every ALI iteration is not
// repeated.  Interpolation points (set_data) are not stored; the solver
// expands them from the stored steps with the same FP64 arithmetic
// (mgr_ray.hpp).
//
// Compiled with -ffp-contract=off.
#include <algorithm>
#include <cstring>
#include <numeric>

#include "mgr_common.hpp"
#include "mgr_ray.hpp"

using namespace mgr;

namespace {

constexpr int32_t ERR_NO_NEIGHBOUR = 1;
constexpr int32_t ERR_TOO_LONG     = 8;
constexpr int32_t ERR_TOO_MANY     = 16;   // a step needs more than 10000 points

constexpr int64_t TRACE_RECORD_MAX   = 4096;       // recorded cells per walk, at most
constexpr int64_t TRACE_RECORD_BYTES = 64LL << 30;  // record buffer, at most

// positive doubles order like their bit patterns
__device__ __forceinline__ void atomic_minmax_pos_double(uint64_t *addr, double lo, double hi)
{
    if (lo > 0.0) atomicMin((unsigned long long *)addr, (unsigned long long)__double_as_longlong(lo));
    if (hi > 0.0) atomicMax((unsigned long long *)addr + 1, (unsigned long long)__double_as_longlong(hi));
}

// Cells recorded by the counting pass: the next cell of every step, four to
// an int4, up to `cap` steps per (origin, ray).  The filling pass replays
// them (replay_walk) instead of scanning the neighbours again.  The step
// index j runs in step on every active lane of a wave, so the select below
// branches on a uniform value.
__device__ __forceinline__ void put_cell(int4 &acc, int j, int32_t c)
{
    switch (j & 3) {
    case 0: acc.x = c; break;
    case 1: acc.y = c; break;
    case 2: acc.z = c; break;
    default: acc.w = c; break;
    }
}

// The walk of one (origin, ray): Geometry::trace_ray (geometry.tpp:21-95)
// with set_data's interpolation split of every step (geometry.tpp:100-207).
// The ray advances along (nx, ny, nz); shifts are co-moving,
// 1 - (v_c - v_o).n, or for REST the rest-frame 1 - v_c.s with s the
// direction of get_doppler_shift<Rest>.  FILL writes the steps from w on;
// otherwise, with rec non-null, the walk records its first `cap` cells.
struct Walk {
    int32_t steps, points, last_cell;   // last_cell: cell of the last emitted point, -1: none
    double  smin, smax, last_shift;     // shift range of the walk (origin included)
};

// Candidate masks of the walk's ray r: entry c Rmax + r holds the bits of
// cell c's candidate neighbours and, in bit 63, whether c is a boundary cell
// (read when the walk is MASKED).
constexpr uint64_t FWD_BOUNDARY = 1ull << 63;
struct Cands {
    const uint64_t *fwd;
    int Rmax, r;
    __device__ __forceinline__ uint64_t of(int32_t c) const { return fwd[(int64_t)c * Rmax + r]; }
};

template <bool FILL, bool REST, bool MASKED>
__device__ __forceinline__ Walk trace_walk(int64_t N, const double *__restrict__ pos,
                                           const double *__restrict__ vel, const int64_t *__restrict__ nb_off,
                                           const NbRec *__restrict__ nb, Cands cd,
                                           const uint8_t *__restrict__ is_boundary, int32_t o,
                                           double nx, double ny, double nz, double sx, double sy,
                                           double sz, double dsm, int64_t w, int4 *__restrict__ rec, int cap,
                                           int32_t *__restrict__ seg_cell,
                                           double *__restrict__ seg_shift, double *__restrict__ seg_dZ,
                                           int4 *__restrict__ seg_f32, float *__restrict__ seg_hdz32,
                                           int32_t *__restrict__ err_flag)
{
    Walk k;
    k.steps = 0; k.points = 0; k.last_cell = -1;
    k.smin = 1.0; k.smax = 1.0; k.last_shift = 1.0;
    double Z = 0.0, dZ = 0.0;
    const double ox = pos[3 * (int64_t)o + 0], oy = pos[3 * (int64_t)o + 1], oz = pos[3 * (int64_t)o + 2];
    int4 acc = make_int4(0, 0, 0, 0);
    int32_t nxt = MASKED ? get_next_masked(nb + nb_off[o], cd.of(o) & ~FWD_BOUNDARY, ox, oy, oz, nx, ny, nz, Z, dZ)
                         : get_next(nb, nb_off, ox, oy, oz, nx, ny, nz, o, Z, dZ);
    if (nxt == -1) return k;
    if (!FILL && rec) put_cell(acc, 0, nxt);
    // what the next step needs of nxt, loaded together: its candidates (and
    // boundary bit) and its neighbour list
    uint64_t fn = MASKED ? cd.of(nxt) : 0;
    const NbRec *ln = MASKED ? nb + nb_off[nxt] : nullptr;
    double shift_crt = REST ? rest_shift(vel, o, sx, sy, sz) : doppler_shift(vel, o, o, nx, ny, nz);
    double shift_nxt = REST ? rest_shift(vel, nxt, sx, sy, sz) : doppler_shift(vel, o, nxt, nx, ny, nz);
    k.smin = fmin(k.smin, shift_crt);
    k.smax = fmax(k.smax, shift_crt);
    for (;;) {
        const StepSplit sp = split_step(shift_crt, shift_nxt, dZ, dsm);
        // get_required_npoints throws for more than 10000 points
        // (geometry.tpp:484), before any radiation field is computed
        if (!FILL && sp.interp && sp.n == 0) atomicOr(err_flag, ERR_TOO_MANY);
        k.points += sp.n;
        k.smin = fmin(k.smin, shift_nxt);
        k.smax = fmax(k.smax, shift_nxt);
        if (FILL) {
            seg_cell[w] = nxt;
            seg_shift[w] = shift_nxt;
            seg_dZ[w] = dZ;
            // the FP32 kernel's split: sub-point m at shift offset s0 + m ds
            const int half = sp.interp ? sp.half : 0;
            const float s0 = (float)((sp.interp ? shift_crt : shift_nxt) - 1.0);
            const float ds = (float)sp.dshift;
            seg_f32[w] = make_int4(nxt, sp.n | (half << STEP_HALF_SHIFT), __float_as_int(s0), __float_as_int(ds));
            seg_hdz32[w] = (float)(sp.dZ * HDZ_SCALE);
            if (sp.n > 0) {   // sub-point m = n >= half belongs to the next cell
                k.last_cell = nxt;
                k.last_shift = sp.interp ? sub_shift(shift_crt, sp.n, sp.dshift) : shift_nxt;
            }
            w++;
        }
        k.steps++;
        if (MASKED ? (fn & FWD_BOUNDARY) != 0 : is_boundary[nxt] != 0) break;
        if (k.steps >= N) { atomicOr(err_flag, ERR_TOO_LONG); break; }
        shift_crt = shift_nxt;
        nxt = MASKED ? get_next_masked(ln, fn & ~FWD_BOUNDARY, ox, oy, oz, nx, ny, nz, Z, dZ)
                     : get_next(nb, nb_off, ox, oy, oz, nx, ny, nz, nxt, Z, dZ);
        if (nxt < 0) { atomicOr(err_flag, ERR_NO_NEIGHBOUR); break; }
        if (MASKED) {
            fn = cd.of(nxt);
            ln = nb + nb_off[nxt];
        }
        const int j = k.steps;   // index of this next step
        if (!FILL && rec && j < cap) {
            if ((j & 3) == 0) rec[(j >> 2) - 1] = acc;
            put_cell(acc, j, nxt);
        }
        shift_nxt = REST ? rest_shift(vel, nxt, sx, sy, sz) : doppler_shift(vel, o, nxt, nx, ny, nz);
    }
    // the last, partly filled block of the record
    if (!FILL && rec && k.steps <= cap) rec[(k.steps - 1) >> 2] = acc;
    return k;
}

// The filling pass of a walk whose `ns` cells the counting pass recorded:
// the same steps as trace_walk<true, false> without the neighbour scans.
// Z_new of each step is recomputed from the cell's position (the neighbour
// records hold copies of the same positions) with get_next's operations, so
// dZ, Z and the shifts are bit-identical.  The steps go four at a time, one
// record block each: the four cells' positions and velocities are loaded
// together, and the block is written with 16-byte stores from w (a multiple
// of 4: every walk's steps start on a block), the slots past the walk's
// last step zeroed.
__device__ __forceinline__ Walk replay_walk(const double *__restrict__ pos, const double *__restrict__ vel,
                                            int32_t o, double nx, double ny, double nz, double dsm, int64_t w,
                                            const int4 *__restrict__ rec, int ns,
                                            int32_t *__restrict__ seg_cell, double *__restrict__ seg_shift,
                                            double *__restrict__ seg_dZ, int4 *__restrict__ seg_f32,
                                            float *__restrict__ seg_hdz32)
{
#pragma clang fp contract(off)
    Walk k;
    k.steps = ns; k.points = 0; k.last_cell = -1;
    k.smin = 1.0; k.smax = 1.0; k.last_shift = 1.0;
    const double ox = pos[3 * (int64_t)o + 0], oy = pos[3 * (int64_t)o + 1], oz = pos[3 * (int64_t)o + 2];
    const double vox = vel[3 * (int64_t)o + 0], voy = vel[3 * (int64_t)o + 1], voz = vel[3 * (int64_t)o + 2];
    double Z = 0.0;
    double shift_crt = doppler_shift_v(vox, voy, voz, vox, voy, voz, nx, ny, nz);
    for (int j0 = 0; j0 < ns; j0 += 4) {
        const int4 blk = rec[j0 >> 2];
        const int32_t cb[4] = {blk.x, blk.y, blk.z, blk.w};
        double p[4][3], v[4][3];
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (j0 + u < ns)
                for (int d = 0; d < 3; d++) {
                    p[u][d] = pos[3 * (int64_t)cb[u] + d];
                    v[u][d] = vel[3 * (int64_t)cb[u] + d];
                }
        int32_t c4[4];
        double sh4[4], dz4[4];
        float hz4[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (j0 + u < ns) {
                const double Rx = p[u][0] - ox;
                const double Ry = p[u][1] - oy;
                const double Rz = p[u][2] - oz;
                const double Z_new = dot3(Rx, Ry, Rz, nx, ny, nz);
                const double dZ = Z_new - Z;
                Z = Z + dZ;
                const double shift_nxt = doppler_shift_v(v[u][0], v[u][1], v[u][2], vox, voy, voz, nx, ny, nz);
                const StepSplit sp = split_step(shift_crt, shift_nxt, dZ, dsm);
                k.points += sp.n;
                const int half = sp.interp ? sp.half : 0;
                const float s0 = (float)((sp.interp ? shift_crt : shift_nxt) - 1.0);
                const float ds = (float)sp.dshift;
                seg_f32[w + u] = make_int4(cb[u], sp.n | (half << STEP_HALF_SHIFT), __float_as_int(s0),
                                           __float_as_int(ds));
                c4[u] = cb[u];
                sh4[u] = shift_nxt;
                dz4[u] = dZ;
                hz4[u] = (float)(sp.dZ * HDZ_SCALE);
                if (sp.n > 0) {
                    k.last_cell = cb[u];
                    k.last_shift = sp.interp ? sub_shift(shift_crt, sp.n, sp.dshift) : shift_nxt;
                }
                shift_crt = shift_nxt;
            } else {   // past the last step: an empty slot (no points)
                seg_f32[w + u] = make_int4(0, 0, 0, 0);
                c4[u] = 0;
                sh4[u] = 0.0;
                dz4[u] = 0.0;
                hz4[u] = 0.0f;
            }
        }
        *reinterpret_cast<int4 *>(seg_cell + w) = make_int4(c4[0], c4[1], c4[2], c4[3]);
        reinterpret_cast<double2 *>(seg_shift + w)[0] = make_double2(sh4[0], sh4[1]);
        reinterpret_cast<double2 *>(seg_shift + w)[1] = make_double2(sh4[2], sh4[3]);
        reinterpret_cast<double2 *>(seg_dZ + w)[0] = make_double2(dz4[0], dz4[1]);
        reinterpret_cast<double2 *>(seg_dZ + w)[1] = make_double2(dz4[2], dz4[3]);
        *reinterpret_cast<float4 *>(seg_hdz32 + w) = make_float4(hz4[0], hz4[1], hz4[2], hz4[3]);
        w += 4;
    }
    return k;
}

// One thread per (origin o, ray r < Rmax); r beyond the origin's own ray
// count exits.  Entry q = cell_qbase[o] + r of the (origin, ray) arrays; the
// ray direction is the global ray m_rbase[model] + r.  With `cells` set the
// counting pass records the first `cap` cells of every walk, and the filling
// pass leaves the walks that fit to k_replay.
template <bool FILL, bool MASKED>
__global__ void k_trace(int64_t N, int Rmax, const int32_t *__restrict__ cell_model,
                        const int64_t *__restrict__ cell_qbase, const int32_t *__restrict__ m_R,
                        const int32_t *__restrict__ m_rbase, const double *__restrict__ pos,
                        const double *__restrict__ vel, const int64_t *__restrict__ nb_off,
                        const NbRec *__restrict__ nb, const uint64_t *__restrict__ fwd,
                        const uint8_t *__restrict__ is_boundary,
                        const double *__restrict__ rays, const double *__restrict__ dshift_max,
                        int4 *__restrict__ cells, int cap,
                        int32_t *__restrict__ nseg, int32_t *__restrict__ npts,
                        const int64_t *__restrict__ seg_off, int32_t *__restrict__ seg_cell,
                        double *__restrict__ seg_shift, double *__restrict__ seg_dZ,
                        int4 *__restrict__ seg_f32, float *__restrict__ seg_hdz32,
                        int32_t *__restrict__ far_cell, double *__restrict__ far_shift,
                        uint64_t *__restrict__ shift_bits, int32_t *__restrict__ err_flag)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * Rmax) return;
    const int32_t o = (int32_t)(t / Rmax);
    const int r = (int)(t % Rmax);
    const int32_t mdl = cell_model[o];
    if (r >= m_R[mdl]) return;
    const int64_t q = cell_qbase[o] + r;
    const int64_t rg = (int64_t)m_rbase[mdl] + r;
    const double nx = rays[3 * rg + 0], ny = rays[3 * rg + 1], nz = rays[3 * rg + 2];
    if (FILL && cells && nseg[q] <= cap) return;   // replayed by k_replay
    const Walk k = trace_walk<FILL, false, MASKED>(N, pos, vel, nb_off, nb, Cands{fwd, Rmax, r}, is_boundary,
                                                   o, nx, ny, nz, nx, ny, nz, dshift_max[o], FILL ? seg_off[q] : 0,
                                                   cells ? cells + q * (cap >> 2) : nullptr, cap, seg_cell,
                                                   seg_shift, seg_dZ, seg_f32, seg_hdz32, err_flag);
    if (!FILL) {
        nseg[q] = k.steps;
        npts[q] = k.points;
        atomic_minmax_pos_double(shift_bits, k.smin, k.smax);
    } else {
        far_cell[q] = k.last_cell;
        far_shift[q] = k.last_shift;
        // empty slots up to the next block of four
        for (int64_t e = seg_off[q] + k.steps; e & 3; e++) {
            seg_cell[e] = 0;
            seg_shift[e] = 0.0;
            seg_dZ[e] = 0.0;
            seg_f32[e] = make_int4(0, 0, 0, 0);
            seg_hdz32[e] = 0.0f;
        }
    }
}

// The filling pass of the walks the counting pass recorded (nseg <= cap),
// one thread per (origin, ray) as k_trace.
__global__ void k_replay(int64_t N, int Rmax, const int32_t *__restrict__ cell_model,
                         const int64_t *__restrict__ cell_qbase, const int32_t *__restrict__ m_R,
                         const int32_t *__restrict__ m_rbase, const double *__restrict__ pos,
                         const double *__restrict__ vel, const double *__restrict__ rays,
                         const double *__restrict__ dshift_max, const int4 *__restrict__ cells, int cap,
                         const int32_t *__restrict__ nseg, const int64_t *__restrict__ seg_off,
                         int32_t *__restrict__ seg_cell, double *__restrict__ seg_shift, double *__restrict__ seg_dZ,
                         int4 *__restrict__ seg_f32, float *__restrict__ seg_hdz32, int32_t *__restrict__ far_cell,
                         double *__restrict__ far_shift)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * Rmax) return;
    const int32_t o = (int32_t)(t / Rmax);
    const int r = (int)(t % Rmax);
    const int32_t mdl = cell_model[o];
    if (r >= m_R[mdl]) return;
    const int64_t q = cell_qbase[o] + r;
    const int ns = nseg[q];
    if (ns > cap) return;   // scanned again by k_trace<true>
    const int64_t rg = (int64_t)m_rbase[mdl] + r;
    const double nx = rays[3 * rg + 0], ny = rays[3 * rg + 1], nz = rays[3 * rg + 2];
    const Walk k = replay_walk(pos, vel, o, nx, ny, nz, dshift_max[o], seg_off[q], cells + q * (cap >> 2), ns,
                               seg_cell, seg_shift, seg_dZ, seg_f32, seg_hdz32);
    far_cell[q] = k.last_cell;
    far_shift[q] = k.last_shift;
}

// Rest-frame walks of one model (Geometry::trace_ray<Rest>).  With r_img >= 0
// the pair of an image: entry 2o the antipode of r_img, 2o + 1 the ray r_img
// itself; with r_img < 0 every ray, entry o R + r (for npoints_on_rays<Rest>).
// The rest-frame direction of ray r is r for r < R/2 and its antipode
// otherwise (geometry.tpp:452-460).
template <bool FILL, bool MASKED>
__global__ void k_trace_rest(int64_t N, int R, int r_img, const double *__restrict__ pos,
                             const double *__restrict__ vel, const int64_t *__restrict__ nb_off,
                             const NbRec *__restrict__ nb, const uint64_t *__restrict__ fwd,
                             const uint8_t *__restrict__ is_boundary,
                             const double *__restrict__ rays, const int32_t *__restrict__ antipod,
                             const double *__restrict__ dshift_max, int32_t *__restrict__ nseg,
                             int32_t *__restrict__ npts, const int64_t *__restrict__ seg_off,
                             int32_t *__restrict__ seg_cell, double *__restrict__ seg_shift,
                             double *__restrict__ seg_dZ, int4 *__restrict__ seg_f32,
                             float *__restrict__ seg_hdz32, int32_t *__restrict__ far_cell,
                             double *__restrict__ far_shift, double *__restrict__ origin_shift,
                             uint64_t *__restrict__ shift_bits, int32_t *__restrict__ err_flag)
{
    const int nlist = r_img >= 0 ? 2 : R;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * nlist) return;
    const int32_t o = (int32_t)(t / nlist);
    const int i = (int)(t % nlist);
    const int ray = r_img >= 0 ? (i ? r_img : antipod[r_img]) : i;
    const int rs = ray >= R / 2 ? antipod[ray] : ray;
    const double nx = rays[3 * ray + 0], ny = rays[3 * ray + 1], nz = rays[3 * ray + 2];
    const double sx = rays[3 * rs + 0], sy = rays[3 * rs + 1], sz = rays[3 * rs + 2];
    const Walk k = trace_walk<FILL, true, MASKED>(N, pos, vel, nb_off, nb, Cands{fwd, R, ray}, is_boundary, o,
                                          nx, ny, nz, sx, sy, sz, dshift_max[o], FILL ? seg_off[t] : 0,
                                          nullptr, 0, seg_cell, seg_shift, seg_dZ, seg_f32, seg_hdz32, err_flag);
    if (!FILL) {
        nseg[t] = k.steps;
        npts[t] = k.points;
        if (shift_bits) atomic_minmax_pos_double(shift_bits, k.smin, k.smax);
    } else {
        far_cell[t] = k.last_cell;
        far_shift[t] = k.last_shift;
        origin_shift[t] = rest_shift(vel, o, sx, sy, sz);
    }
}

// Candidate masks of every (cell c, ray r < its model's ray count), entry
// c Rmax + r: bit j is set unless entry j of c's neighbour list lies behind c
// along the ray, (p_j - p_c).n < -tau; bit 63 is set for a boundary cell.
// The walk's test Z_new > Z compares (p_j - o).n with the running Z, which
// equals (p_c - o).n up to a few roundings of numbers no larger than
// |p| + |o|; tau = 1e-9 of the model's extent dwarfs that (some 1e-15 of
// it), so a cleared neighbour never passes the test and the walk's choice is
// unchanged.
__global__ void k_forward_masks(int64_t N, int Rmax, const int32_t *__restrict__ cell_model,
                                const int32_t *__restrict__ m_R, const int32_t *__restrict__ m_rbase,
                                const double *__restrict__ pos, const int64_t *__restrict__ nb_off,
                                const NbRec *__restrict__ nb, const uint8_t *__restrict__ is_boundary,
                                const double *__restrict__ rays, double tau, uint64_t *__restrict__ fwd)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * Rmax) return;
    const int32_t c = (int32_t)(t / Rmax);
    const int r = (int)(t % Rmax);
    const int32_t mdl = cell_model[c];
    if (r >= m_R[mdl]) return;
    const int64_t rg = (int64_t)m_rbase[mdl] + r;
    const double nx = rays[3 * rg + 0], ny = rays[3 * rg + 1], nz = rays[3 * rg + 2];
    const double cx = pos[3 * (int64_t)c + 0], cy = pos[3 * (int64_t)c + 1], cz = pos[3 * (int64_t)c + 2];
    const int64_t b = nb_off[c];
    const int deg = (int)(nb_off[c + 1] - b);
    uint64_t m = is_boundary[c] ? FWD_BOUNDARY : 0;
    for (int j = 0; j < deg; j++) {
        const NbRec e = nb[b + j];
        if (!((e.x - cx) * nx + (e.y - cy) * ny + (e.z - cz) * nz < -tau)) m |= 1ull << j;   // NaN: kept
    }
    fwd[t] = m;
}

// Expands the cached steps of one ray into its projected points (what
// Geometry::trace_ray returns), with the solver's expansion arithmetic.
__global__ void k_expand_ray(int32_t o, int64_t q, const int64_t *__restrict__ seg_off,
                             const int32_t *__restrict__ nseg, const int32_t *__restrict__ seg_cell,
                             const double *__restrict__ seg_shift, const double *__restrict__ seg_dZ,
                             const double *__restrict__ dshift_max, int64_t capacity,
                             int64_t *__restrict__ out_cell, double *__restrict__ out_shift,
                             double *__restrict__ out_dZ, int64_t *__restrict__ n_out)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int64_t s0 = seg_off[q];
    const int ns = nseg[q];
    const double dsm = dshift_max[o];
    int64_t n = 0;
    int32_t crt = o;
    double s_crt = 1.0;
    for (int j = 0; j < ns; j++) {
        const int32_t c = seg_cell[s0 + j];
        const double s = seg_shift[s0 + j];
        const double z = seg_dZ[s0 + j];
        const StepSplit sp = split_step(s_crt, s, z, dsm);
        if (!sp.interp) {
            if (n < capacity) { out_cell[n] = c; out_shift[n] = s; out_dZ[n] = z; }
            n++;
        } else {
            for (int m = 1; m <= sp.n; m++) {
                if (n < capacity) {
                    out_cell[n] = (m < sp.half) ? crt : c;
                    out_shift[n] = sub_shift(s_crt, m, sp.dshift);
                    out_dZ[n] = sp.dZ;
                }
                n++;
            }
        }
        crt = c;
        s_crt = s;
    }
    *n_out = n;
}


// Shift intervals in which some bin of a 64-bin chunk (the bins of lane slots
// 64k .. 64k + 63, lane_bin) can meet a line window.  Bin f meets line l at
// shift s when |s nu_f - nu_l| <= window nu_l pw, i.e. s in
// [(1 - window pw) c, (1 + window pw) c] with c = nu_l / nu_f (window:
// MASK_WINDOW for the FP32 kernel, MASK_WINDOW64 for the FP64 one).  Per
// line, the hull (a, b) = (min c, max c) over the chunk's bins whose c can
// reach the trace's shift range [smin, smax] at the widest width pw_max covers
// all of them; it is widened by the step's own width when tested.  Hulls that
// overlap (the same roots of different lines) merge.  cnt = -1 when a chunk
// needs more than MAX_WIN intervals (never skipped).
constexpr int MAX_WIN = 4;

__global__ void k_chunk_windows(int64_t N, int F, int nchunk, bool root_major, const LineTables *__restrict__ tb,
                                const double *__restrict__ nu, double smin, double smax, double pw_max,
                                double window, double *__restrict__ win, int8_t *__restrict__ cnt)
{
    // one wave per (origin, chunk), lane i on the chunk's bin 64 k + i; the
    // hull of each line is a wave minimum and maximum (exact, so in any order)
    const int64_t t = (int64_t)blockIdx.x * (blockDim.x / WAVE) + __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    if (t >= N * nchunk) return;
    const int64_t o = t / nchunk;
    const int k = (int)(t % nchunk);
    const int i = k * WAVE + (int)(threadIdx.x & (WAVE - 1));
    const bool valid = i < F;
    const double nuv = valid ? nu[o * F + lane_bin(i, tb->L, F, root_major)] : 1.0;
    const double reach = window * pw_max * 1.01;
    double wa[MAX_WIN], wb[MAX_WIN];
    int n = 0;
    for (int l = 0; l < tb->L && n >= 0; l++) {
        const double nl = tb->line[l];
        const double c = nl / nuv;
        const bool in = valid && c * (1.0 - reach) <= smax * (1.0 + 1e-9) && c * (1.0 + reach) >= smin * (1.0 - 1e-9);
        if (!__any(in)) continue;
        double a = in ? c : DBL_MAX, b = in ? c : -DBL_MAX;
        for (int off = WAVE / 2; off >= 1; off >>= 1) {
            a = fmin(a, __shfl_xor(a, off));
            b = fmax(b, __shfl_xor(b, off));
        }
        // overlapping intervals (the same roots of different lines) merge
        bool merged = false;
        for (int j = 0; j < n && !merged; j++) {
            if (a <= wb[j] && b >= wa[j]) {
                wa[j] = fmin(wa[j], a);
                wb[j] = fmax(wb[j], b);
                merged = true;
            }
        }
        if (merged) continue;
        if (n == MAX_WIN) { n = -1; break; }
        wa[n] = a;
        wb[n] = b;
        n++;
    }
    if ((threadIdx.x & (WAVE - 1)) != 0) return;
    for (int j = 0; j < n; j++) {
        win[((o * STEP_MASK_CHUNKS + k) * MAX_WIN + j) * 2 + 0] = wa[j];
        win[((o * STEP_MASK_CHUNKS + k) * MAX_WIN + j) * 2 + 1] = wb[j];
    }
    cnt[o * STEP_MASK_CHUNKS + k] = (int8_t)n;
}

// Out-of-window bits of every cached step.  A resolved step has one point, at
// its own shift in its own cell; the sub-points of an interpolated step lie
// between the shifts of its two ends and in its two cells.  Bit k is set when
// that shift range [lo, hi] misses every interval of chunk k widened by the
// (wider) cell width.  One thread per (origin, ray); every ray's steps start
// on a block of four, so a block's cells and shifts are read with 16-byte
// loads and its four mask bytes written as one word (zero past the last
// step).
__global__ void k_step_masks(int64_t N, int Rmax, int nchunk, const int32_t *__restrict__ cell_model,
                             const int64_t *__restrict__ cell_qbase, const int32_t *__restrict__ m_R,
                             const double *__restrict__ pw, const double *__restrict__ dshift_max,
                             const int64_t *__restrict__ seg_off, const int32_t *__restrict__ nseg,
                             const int32_t *__restrict__ seg_cell, const double *__restrict__ seg_shift,
                             double window, const double *__restrict__ win,
                             const int8_t *__restrict__ cnt, uint8_t *__restrict__ mask)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * Rmax) return;
    const int32_t o = (int32_t)(t / Rmax);
    const int r = (int)(t % Rmax);
    if (r >= m_R[cell_model[o]]) return;
    const int64_t q = cell_qbase[o] + r;
    const int64_t s0 = seg_off[q];
    const int ns = nseg[q];
    const double dsm = dshift_max[o];
    int cn[STEP_MASK_CHUNKS];
    for (int k = 0; k < nchunk; k++) cn[k] = cnt[(int64_t)o * STEP_MASK_CHUNKS + k];
    const double *wo = win + (int64_t)o * STEP_MASK_CHUNKS * MAX_WIN * 2;
    int32_t c_prev = o;
    double s_prev = 1.0;   // co-moving shift of the origin
    for (int j0 = 0; j0 < ns; j0 += 4) {
        const int4 cb = *reinterpret_cast<const int4 *>(seg_cell + s0 + j0);
        const double2 sa = reinterpret_cast<const double2 *>(seg_shift + s0 + j0)[0];
        const double2 sb = reinterpret_cast<const double2 *>(seg_shift + s0 + j0)[1];
        const int32_t cs[4] = {cb.x, cb.y, cb.z, cb.w};
        const double ss[4] = {sa.x, sa.y, sb.x, sb.y};
        unsigned word = 0;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (j0 + u >= ns) break;
            const int32_t c = cs[u];
            const double s = ss[u];
            const bool interp = fabs(s - s_prev) > dsm;   // as split_step
            const double lo = interp ? fmin(s_prev, s) : s, hi = interp ? fmax(s_prev, s) : s;
            const double w = window * (interp ? fmax(pw[c_prev], pw[c]) : pw[c]);
            unsigned bits = 0;
            for (int k = 0; k < nchunk; k++) {
                const int n = cn[k];
                if (n < 0) continue;
                bool out = true;
                for (int i = 0; i < n; i++) {
                    const double a = wo[(k * MAX_WIN + i) * 2 + 0];
                    const double b = wo[(k * MAX_WIN + i) * 2 + 1];
                    if (hi >= (1.0 - w) * a && lo <= (1.0 + w) * b) out = false;
                }
                if (out) bits |= 1u << k;
            }
            word |= bits << (8 * u);
            c_prev = c;
            s_prev = s;
        }
        *reinterpret_cast<uint32_t *>(mask + s0 + j0) = word;
    }
}

// Sums over the cached steps of (points x chunks) and of (points x chunks
// whose out-of-window bit is set): one thread per step, a block sum, one
// atomic add per block.
__global__ void k_mask_stats(int64_t T, int nchunk, const int4 *__restrict__ seg_f32,
                             const uint8_t *__restrict__ mask, unsigned long long *__restrict__ out)
{
    __shared__ unsigned long long s_all[256], s_skip[256];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long all = 0, skip = 0;
    if (t < T) {
        const unsigned n = (unsigned)(seg_f32[t].y & STEP_N_MASK);
        all = (unsigned long long)n * nchunk;
        skip = (unsigned long long)n * __popc((unsigned)mask[t] & ((1u << nchunk) - 1u));
    }
    s_all[threadIdx.x] = all;
    s_skip[threadIdx.x] = skip;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s_all[threadIdx.x] += s_all[threadIdx.x + w];
            s_skip[threadIdx.x] += s_skip[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(&out[0], s_all[0]);
        atomicAdd(&out[1], s_skip[0]);
    }
}

// Largest dz' = dZ HDZ_SCALE of the cached steps: non-negative floats order
// as their bit patterns, so a block maximum of the bits and one atomicMax per
// block.
__global__ void k_max_hdz(int64_t T, const float *__restrict__ hdz, unsigned *__restrict__ out)
{
    __shared__ unsigned s_max[256];
    unsigned v = 0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < T; t += (int64_t)gridDim.x * blockDim.x)
        v = max(v, __float_as_uint(hdz[t]));
    s_max[threadIdx.x] = v;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s_max[threadIdx.x] = max(s_max[threadIdx.x], s_max[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, s_max[0]);
}

// Cell of sub-point i (0: nearest the origin) of entry q: the steps' split
// (Geometry::set_data, geometry.tpp:100-207) assigns sub-points m < half of a
// step to the step's current cell and the others to its next cell.
__device__ int32_t sub_point_cell(const int64_t *__restrict__ seg_off, const int32_t *__restrict__ nseg,
                                  const int32_t *__restrict__ seg_cell, const double *__restrict__ seg_shift,
                                  const double *__restrict__ seg_dZ, int64_t q, int32_t o, double dsm, int i)
{
    const int64_t b = seg_off[q];
    const int ns = nseg[q];
    for (int j = 0; j < ns; j++) {
        const double sc = j > 0 ? seg_shift[b + j - 1] : 1.0;
        const StepSplit sp = split_step(sc, seg_shift[b + j], seg_dZ[b + j], dsm);
        if (i < sp.n) return (i + 1 < (sp.interp ? sp.half : 0)) ? (j > 0 ? seg_cell[b + j - 1] : o) : seg_cell[b + j];
        i -= sp.n;
    }
    return -1;
}

// Cells of the Lambda* band elements of every (origin, ray pair), one model
// (RayPair::update_Lambda, raypair_lambda.tpp:129-146).  The longer ray of
// the pair is the low-index side (sim_radiation.tpp:164-171), ties to the
// antipode.  Band m's lower element is at point n_ar - m, sub-point m-1 of the
// longer ray; it exists when the shorter ray has a point, so that the pair's
// solve wrote L_lower[m][n_ar-m] (raypair_solver.tpp:381-388).  Its upper
// element is at point n_ar + m, sub-point m-1 of the shorter ray; it exists
// when that ray has more than 2m points (L_upper[m][n_ar+m] written).  Other
// slots get -1.
__global__ void k_band_cells(int64_t N, int Rh, int nband, const int64_t *__restrict__ cell_qbase,
                             const int32_t *__restrict__ antipod, const int64_t *__restrict__ seg_off,
                             const int32_t *__restrict__ nseg, const int32_t *__restrict__ npts,
                             const int32_t *__restrict__ seg_cell, const double *__restrict__ seg_shift,
                             const double *__restrict__ seg_dZ, const double *__restrict__ dshift_max,
                             int32_t *__restrict__ band_cell)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * Rh) return;
    const int32_t o = (int32_t)(t / Rh);
    const int r = (int)(t - (int64_t)o * Rh);
    const int64_t qr = cell_qbase[o] + r, qa = cell_qbase[o] + antipod[r];
    const int n_r = npts[qr], n_a = npts[qa];
    const bool r_longer = n_r > n_a;
    const int64_t qA = r_longer ? qr : qa, qB = r_longer ? qa : qr;
    const int nA = r_longer ? n_r : n_a, nB = r_longer ? n_a : n_r;
    const double dsm = dshift_max[o];
    int32_t *out = band_cell + t * 2 * nband;
    for (int m = 1; m <= nband; m++) {
        out[m - 1] = (nB >= 1 && nA >= m)
                         ? sub_point_cell(seg_off, nseg, seg_cell, seg_shift, seg_dZ, qA, o, dsm, m - 1) : -1;
        out[nband + m - 1] = (nB > 2 * m)
                         ? sub_point_cell(seg_off, nseg, seg_cell, seg_shift, seg_dZ, qB, o, dsm, m - 1) : -1;
    }
}

}  // namespace

namespace mgr {

// Band cells of the current trace and, on the host, their reverse map: for
// every cell the slots (origin S + slot) that name it, in ascending order.
int launch_band_cells(mgr_ctx *ctx)
{
    const int64_t N = ctx->N, S = ctx->band_S;
    const int Rh = (int)(ctx->R / 2);
    hipLaunchKernelGGL(k_band_cells, dim3((unsigned)((N * Rh + 127) / 128)), dim3(128), 0, ctx->stream, N, Rh,
                       ctx->nband, ctx->d_cell_qbase, ctx->d_antipod, ctx->seg_off, ctx->nseg, ctx->npts,
                       ctx->seg_cell, ctx->seg_shift, ctx->seg_dZ, ctx->dshift_max, ctx->band_cell);
    MGR_LAUNCH_CHECK();
    std::vector<int32_t> cell(N * S);
    MGR_HIP(hipStreamSynchronize(ctx->stream));
    MGR_HIP(hipMemcpy(cell.data(), ctx->band_cell, N * S * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<int64_t> off(N + 1, 0);
    for (int64_t i = 0; i < N * S; i++)
        if (cell[i] >= 0) off[cell[i] + 1]++;
    for (int64_t c = 0; c < N; c++) off[c + 1] += off[c];
    std::vector<int64_t> rev(std::max<int64_t>(off[N], 1)), pos(off.begin(), off.end() - 1);
    for (int64_t i = 0; i < N * S; i++)
        if (cell[i] >= 0) rev[pos[cell[i]]++] = i;
    if (ctx->band_rev_off) (void)hipFree(ctx->band_rev_off);
    if (ctx->band_rev) (void)hipFree(ctx->band_rev);
    ctx->band_rev_off = nullptr;
    ctx->band_rev = nullptr;
    MGR_HIP(hipMalloc((void **)&ctx->band_rev_off, (N + 1) * sizeof(int64_t)));
    MGR_HIP(hipMalloc((void **)&ctx->band_rev, rev.size() * sizeof(int64_t)));
    MGR_HIP(hipMemcpy(ctx->band_rev_off, off.data(), (N + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    MGR_HIP(hipMemcpy(ctx->band_rev, rev.data(), rev.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    return MGR_OK;
}

// Device memory held for the length of a host function.
struct DeviceBuffer {
    void *p = nullptr;
    void release() { if (p) (void)hipFree(p); p = nullptr; }
    ~DeviceBuffer() { release(); }
};

static int trace_error(mgr_ctx *ctx, int32_t errv)
{
    if (errv & ERR_TOO_MANY) return ctx->fail(MGR_ERR_TRACE, "Too many (> 10 000) interpolations needed!");
    return ctx->fail(MGR_ERR_TRACE, errv & ERR_NO_NEIGHBOUR ? "a ray found no forward neighbour inside the mesh"
                                                            : "a ray visited more steps than there are cells");
}

// Candidate masks of every (cell, ray), once per context (the geometry is
// static); none when a neighbour list is longer than 63.
int launch_forward_masks(mgr_ctx *ctx)
{
    const int64_t n = ctx->N * ctx->R;
    if (ctx->fwd_mask || !ctx->trace_masks || ctx->nb_kmax > 63 || n == 0) return MGR_OK;
    MGR_HIP(hipMalloc((void **)&ctx->fwd_mask, n * sizeof(uint64_t)));
    hipLaunchKernelGGL(k_forward_masks, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, ctx->stream, ctx->N,
                       (int)ctx->R, ctx->d_cell_model, ctx->d_m_R, ctx->d_m_rbase, ctx->pos, ctx->nb_off, ctx->nb_rec,
                       ctx->is_boundary, ctx->rays, ctx->fwd_tau, ctx->fwd_mask);
    MGR_LAUNCH_CHECK();
    return MGR_OK;
}

int launch_trace(mgr_ctx *ctx)
{
    const int64_t N = ctx->N, R = ctx->R, NR = ctx->NQ;
    const unsigned grid = (unsigned)((N * R + 127) / 128);
    if (!ctx->nseg) {
        MGR_HIP(hipMalloc((void **)&ctx->nseg, NR * sizeof(int32_t)));
        MGR_HIP(hipMalloc((void **)&ctx->npts, NR * sizeof(int32_t)));
        MGR_HIP(hipMalloc((void **)&ctx->seg_off, (NR + 1) * sizeof(int64_t)));
    }
    static const uint64_t one_bits[2] = {0x3FF0000000000000ull, 0x3FF0000000000000ull};   // 1.0, 1.0
    MGR_HIP(hipMemcpyAsync(ctx->shift_bits, one_bits, 2 * sizeof(uint64_t), hipMemcpyHostToDevice, ctx->stream));
    MGR_HIP(hipMemsetAsync(ctx->err_flag, 0, sizeof(int32_t), ctx->stream));
    // The counting pass records the cells of every walk of up to `cap` steps
    // (a multiple of 4) in a buffer of at most a quarter of the free memory;
    // the filling pass replays them.  Longer walks are scanned again.
    DeviceBuffer cells;
    int cap = 0;
    if (NR > 0) {
        size_t freeb = 0, totb = 0;
        if (hipMemGetInfo(&freeb, &totb) == hipSuccess) {
            const int64_t budget = std::min<int64_t>((int64_t)(freeb / 4), TRACE_RECORD_BYTES);
            cap = (int)std::min<int64_t>(ctx->trace_record_steps, budget / NR / (int64_t)sizeof(int32_t) / 4 * 4);
            if (cap < 4 || hipMalloc(&cells.p, (size_t)NR * cap * sizeof(int32_t)) != hipSuccess) {
                (void)hipGetLastError();
                cells.p = nullptr;
                cap = 0;
            }
        }
    }
    int rc = launch_forward_masks(ctx);
    if (rc) return rc;
    auto count = ctx->fwd_mask ? k_trace<false, true> : k_trace<false, false>;
    hipLaunchKernelGGL(count, dim3(grid), dim3(128), 0, ctx->stream, N, (int)R, ctx->d_cell_model,
                       ctx->d_cell_qbase, ctx->d_m_R, ctx->d_m_rbase, ctx->pos, ctx->vel, ctx->nb_off, ctx->nb_rec,
                       ctx->fwd_mask, ctx->is_boundary, ctx->rays, ctx->dshift_max, (int4 *)cells.p, cap, ctx->nseg,
                       ctx->npts, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                       ctx->shift_bits, ctx->err_flag);
    MGR_LAUNCH_CHECK();
    std::vector<int32_t> ns(NR), np(NR);
    int32_t errv = 0;
    uint64_t sbits[2] = {0, 0};
    MGR_HIP(hipStreamSynchronize(ctx->stream));
    MGR_HIP(hipMemcpy(&errv, ctx->err_flag, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (errv) return trace_error(ctx, errv);
    MGR_HIP(hipMemcpy(ns.data(), ctx->nseg, NR * sizeof(int32_t), hipMemcpyDeviceToHost));
    MGR_HIP(hipMemcpy(np.data(), ctx->npts, NR * sizeof(int32_t), hipMemcpyDeviceToHost));
    MGR_HIP(hipMemcpy(sbits, ctx->shift_bits, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost));
    std::vector<int64_t> off(NR + 1, 0);
    int64_t max_steps = 0;
    for (int64_t q = 0; q < NR; q++) {   // every walk starts on a block of four steps
        off[q + 1] = off[q] + (ns[q] + 3) / 4 * 4;
        max_steps = std::max<int64_t>(max_steps, ns[q]);
    }
    ctx->total_segments = off[NR];
    double smin, smax;
    std::memcpy(&smin, &sbits[0], sizeof(double));
    std::memcpy(&smax, &sbits[1], sizeof(double));
    ctx->shift_min = std::min(1.0, smin);
    ctx->shift_max = std::max(1.0, smax);

    // ray-pair statistics: ndep = n_ar + n_r + 1 (sim_radiation.tpp:164-178)
    int64_t sum_ndep = 0, solved = 0;
    for (int64_t k = 0; k < ctx->M; k++) {
        const int64_t Rk = ctx->m_R[k], rb = ctx->m_rbase[k];
        for (int64_t o = ctx->m_cell_off[k]; o < ctx->m_cell_off[k + 1]; o++) {
            const int64_t qb = ctx->h_cell_qbase[o];
            for (int64_t r = 0; r < Rk / 2; r++) {
                const int64_t nd = (int64_t)np[qb + r] + np[qb + ctx->antipod[rb + r] - rb] + 1;
                sum_ndep += nd;
                solved += (nd > 1);
            }
        }
    }
    ctx->sum_ndep = sum_ndep;
    ctx->n_pairs_solved = solved;

    // the ALI graphs captured pointers into the old cache
    ctx->graph_precision = -1;
    ctx->masks_valid = false;
    ctx->masks64_valid = false;
    ctx->far_invalidate();
    if (ctx->seg_cell) {
        (void)hipFree(ctx->seg_cell); (void)hipFree(ctx->seg_shift); (void)hipFree(ctx->seg_dZ);
        (void)hipFree(ctx->seg_f32); (void)hipFree(ctx->seg_hdz32); (void)hipFree(ctx->step_mask);
        if (ctx->step_mask64) (void)hipFree(ctx->step_mask64);
        ctx->step_mask64 = nullptr;
        ctx->seg_cell = nullptr; ctx->seg_shift = nullptr; ctx->seg_dZ = nullptr;
        ctx->seg_f32 = nullptr; ctx->seg_hdz32 = nullptr; ctx->step_mask = nullptr;
    }
    const int64_t T = std::max<int64_t>(ctx->total_segments, 1);
    MGR_HIP(hipMalloc((void **)&ctx->step_mask, T * sizeof(uint8_t)));
    MGR_HIP(hipMalloc((void **)&ctx->seg_cell, T * sizeof(int32_t)));
    MGR_HIP(hipMalloc((void **)&ctx->seg_shift, T * sizeof(double)));
    MGR_HIP(hipMalloc((void **)&ctx->seg_dZ, T * sizeof(double)));
    MGR_HIP(hipMalloc((void **)&ctx->seg_f32, T * sizeof(int4)));
    MGR_HIP(hipMalloc((void **)&ctx->seg_hdz32, T * sizeof(float)));
    if (!ctx->far_cell) {
        MGR_HIP(hipMalloc((void **)&ctx->far_cell, std::max<int64_t>(NR, 1) * sizeof(int32_t)));
        MGR_HIP(hipMalloc((void **)&ctx->far_shift, std::max<int64_t>(NR, 1) * sizeof(double)));
    }
    MGR_HIP(hipMemcpy(ctx->seg_off, off.data(), (NR + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    if (cells.p)
        hipLaunchKernelGGL(k_replay, dim3(grid), dim3(128), 0, ctx->stream, N, (int)R, ctx->d_cell_model,
                           ctx->d_cell_qbase, ctx->d_m_R, ctx->d_m_rbase, ctx->pos, ctx->vel, ctx->rays,
                           ctx->dshift_max, (const int4 *)cells.p, cap, ctx->nseg, ctx->seg_off, ctx->seg_cell,
                           ctx->seg_shift, ctx->seg_dZ, ctx->seg_f32, ctx->seg_hdz32, ctx->far_cell, ctx->far_shift);
    if (!cells.p || max_steps > cap) {   // the walks not recorded
        auto fill = ctx->fwd_mask ? k_trace<true, true> : k_trace<true, false>;
        hipLaunchKernelGGL(fill, dim3(grid), dim3(128), 0, ctx->stream, N, (int)R, ctx->d_cell_model,
                           ctx->d_cell_qbase, ctx->d_m_R, ctx->d_m_rbase, ctx->pos, ctx->vel, ctx->nb_off, ctx->nb_rec,
                           ctx->fwd_mask, ctx->is_boundary, ctx->rays, ctx->dshift_max, (int4 *)cells.p, cap,
                           ctx->nseg, nullptr, ctx->seg_off, ctx->seg_cell, ctx->seg_shift, ctx->seg_dZ,
                           ctx->seg_f32, ctx->seg_hdz32, ctx->far_cell, ctx->far_shift, nullptr, ctx->err_flag);
    }
    MGR_LAUNCH_CHECK();
    MGR_HIP(hipStreamSynchronize(ctx->stream));
    cells.release();
    if (ctx->nband > 0) {
        const int rc = launch_band_cells(ctx);
        if (rc) return rc;
    }
    ctx->traced = true;
    return MGR_OK;
}

// Computes the out-of-window bits of the cached steps for the current
// frequencies (once per trace and spectral discretisation), for the FP32
// kernel or (fp64) the FP64 one.  Models with more than STEP_MASK_CHUNKS
// chunks of 64 bins get bits for the first ones only.
int launch_step_masks(mgr_ctx *ctx, bool fp64)
{
    bool &valid = fp64 ? ctx->masks64_valid : ctx->masks_valid;
    if (valid || !ctx->traced || !ctx->spectral_set) return MGR_OK;
    uint8_t *&bits_out = fp64 ? ctx->step_mask64 : ctx->step_mask;
    if (!bits_out) MGR_HIP(hipMalloc((void **)&bits_out, std::max<int64_t>(ctx->total_segments, 1)));
    const double window = fp64 ? MASK_WINDOW64 : MASK_WINDOW;
    const int64_t N = ctx->N;
    const int nchunk = (int)std::min<int64_t>((ctx->F + WAVE - 1) / WAVE, STEP_MASK_CHUNKS);
    // Root-major lanes when the lines are further apart than the trace's
    // shift range plus both bins' spans and windows: then no bin meets a
    // second line and each wave's bins leave their windows together.
    bool rm = false;
    const mgr::LineTables &tb = ctx->h_tables;
    if (!ctx->spectral_image && tb.L > 1 && ctx->F == (int64_t)tb.L * tb.Q) {
        const double need = (ctx->shift_max - ctx->shift_min) + 2.0 * ctx->pw_max * (ctx->root_max + MASK_WINDOW);
        rm = true;
        for (int l = 0; l + 1 < tb.L; l++)
            if (!(tb.line[l + 1] > tb.line[l] * (1.0 + 2.0 * need))) rm = false;
    }
    if (rm != ctx->root_major) ctx->graph_precision = -1;   // captured launches hold the old order
    ctx->root_major = rm;
    double *win = nullptr;
    int8_t *cnt = nullptr;
    MGR_HIP(hipMalloc((void **)&win, (size_t)N * STEP_MASK_CHUNKS * MAX_WIN * 2 * sizeof(double)));
    if (hipMalloc((void **)&cnt, (size_t)N * STEP_MASK_CHUNKS) != hipSuccess) {
        (void)hipFree(win);
        return ctx->fail(MGR_ERR_DEVICE, "out of device memory");
    }
    hipLaunchKernelGGL(k_chunk_windows, dim3((unsigned)((N * nchunk + 3) / 4)), dim3(4 * WAVE), 0, ctx->stream, N,
                       (int)ctx->F, nchunk, rm, ctx->tables, ctx->nu, ctx->shift_min, ctx->shift_max, ctx->pw_max,
                       window, win, cnt);
    hipLaunchKernelGGL(k_step_masks, dim3((unsigned)((N * ctx->R + 127) / 128)), dim3(128), 0, ctx->stream, N,
                       (int)ctx->R, nchunk, ctx->d_cell_model, ctx->d_cell_qbase, ctx->d_m_R, ctx->pw_cmax,
                       ctx->dshift_max, ctx->seg_off,
                       ctx->nseg, ctx->seg_cell, ctx->seg_shift, window, win, cnt, bits_out);
    // the largest dz' bounds the row closings of out-of-window sub-points (FP32)
    unsigned *dmax = reinterpret_cast<unsigned *>(cnt);   // N * STEP_MASK_CHUNKS >= 4 bytes, no longer read
    hipError_t le = hipGetLastError();
    if (le == hipSuccess) le = hipMemsetAsync(dmax, 0, sizeof(unsigned), ctx->stream);
    if (le == hipSuccess && ctx->total_segments > 0 && !fp64) {
        hipLaunchKernelGGL(k_max_hdz, dim3(1024), dim3(256), 0, ctx->stream, ctx->total_segments, ctx->seg_hdz32,
                           dmax);
        le = hipGetLastError();
    }
    unsigned bits = 0;
    const hipError_t se = hipStreamSynchronize(ctx->stream);
    if (le == hipSuccess && se == hipSuccess) le = hipMemcpy(&bits, dmax, sizeof(unsigned), hipMemcpyDeviceToHost);
    (void)hipFree(win);
    (void)hipFree(cnt);
    if (le != hipSuccess || se != hipSuccess) return ctx->fail(MGR_ERR_DEVICE, "step-mask kernels failed");
    if (!fp64) std::memcpy(&ctx->hdz32_max, &bits, sizeof(float));
    valid = true;
    return MGR_OK;
}

}  // namespace mgr

// Point-chunks of the cached trace (interpolation sub-points x 64-bin chunks
// with a bit, up to STEP_MASK_CHUNKS) and those the FP32 kernel skips.
extern "C" int mgr_window_skip_stats(mgr_ctx *ctx, int64_t *point_chunks, int64_t *skipped)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    if (!ctx->traced) return ctx->fail(MGR_ERR_STATE, "mgr_trace first");
    if (!ctx->spectral_set) return ctx->fail(MGR_ERR_STATE, "compute_spectral_discretisation first");
    MGR_HIP(hipSetDevice(ctx->device));
    int rc = mgr::launch_step_masks(ctx, false);
    if (rc) return rc;
    const int64_t T = ctx->total_segments;
    const int nchunk = (int)std::min<int64_t>((ctx->F + mgr::WAVE - 1) / mgr::WAVE, mgr::STEP_MASK_CHUNKS);
    unsigned long long *d = nullptr, h[2] = {0, 0};
    MGR_HIP(hipMalloc((void **)&d, sizeof(h)));
    hipError_t e = hipMemsetAsync(d, 0, sizeof(h), ctx->stream);
    if (e == hipSuccess && T > 0) {
        hipLaunchKernelGGL(k_mask_stats, dim3((unsigned)((T + 255) / 256)), dim3(256), 0, ctx->stream, T, nchunk,
                           ctx->seg_f32, ctx->step_mask, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return ctx->fail(MGR_ERR_DEVICE, std::string("window-skip stats: ") + hipGetErrorString(e));
    if (point_chunks) *point_chunks = (int64_t)h[0];
    if (skipped) *skipped = (int64_t)h[1];
    return MGR_OK;
}

namespace mgr {

void free_image(mgr_ctx *ctx)
{
    mgr_ctx::ImageTrace &g = ctx->img;
    for (void *p : {(void *)g.seg_off, (void *)g.nseg, (void *)g.npts, (void *)g.seg_cell, (void *)g.seg_shift,
                    (void *)g.seg_dZ, (void *)g.seg_f32, (void *)g.seg_hdz32, (void *)g.far_cell,
                    (void *)g.far_shift, (void *)g.origin_shift, (void *)g.I_m, (void *)g.I_p})
        if (p) (void)hipFree(p);
    g = mgr_ctx::ImageTrace();
}


// Rest-frame trace of ray `ray` and its antipode for an image: counting
// pass, host scan, filling pass into ctx->img (kept until another ray is
// imaged; the geometry is static).
int launch_trace_image(mgr_ctx *ctx, int64_t ray)
{
    if (ctx->img.ray == ray) return MGR_OK;
    const int64_t N = ctx->N, R = ctx->R, NQ = 2 * N;
    mgr_ctx::ImageTrace &g = ctx->img;
    g.ray = -1;
    const int mrc = launch_forward_masks(ctx);
    if (mrc) return mrc;
    if (!g.nseg) {
        MGR_HIP(hipMalloc((void **)&g.nseg, NQ * sizeof(int32_t)));
        MGR_HIP(hipMalloc((void **)&g.npts, NQ * sizeof(int32_t)));
        MGR_HIP(hipMalloc((void **)&g.seg_off, (NQ + 1) * sizeof(int64_t)));
        MGR_HIP(hipMalloc((void **)&g.far_cell, NQ * sizeof(int32_t)));
        MGR_HIP(hipMalloc((void **)&g.far_shift, NQ * sizeof(double)));
        MGR_HIP(hipMalloc((void **)&g.origin_shift, NQ * sizeof(double)));
    }
    const unsigned grid = (unsigned)((NQ + 127) / 128);
    static const uint64_t one_bits[2] = {0x3FF0000000000000ull, 0x3FF0000000000000ull};   // 1.0, 1.0
    MGR_HIP(hipMemcpyAsync(ctx->shift_bits, one_bits, 2 * sizeof(uint64_t), hipMemcpyHostToDevice, ctx->stream));
    MGR_HIP(hipMemsetAsync(ctx->err_flag, 0, sizeof(int32_t), ctx->stream));
    auto count = ctx->fwd_mask ? k_trace_rest<false, true> : k_trace_rest<false, false>;
    hipLaunchKernelGGL(count, dim3(grid), dim3(128), 0, ctx->stream, N, (int)R, (int)ray, ctx->pos, ctx->vel,
                       ctx->nb_off, ctx->nb_rec, ctx->fwd_mask, ctx->is_boundary, ctx->rays,
                       ctx->d_antipod, ctx->dshift_max, g.nseg, g.npts, nullptr, nullptr, nullptr, nullptr, nullptr,
                       nullptr, nullptr, nullptr, nullptr, ctx->shift_bits, ctx->err_flag);
    MGR_LAUNCH_CHECK();
    MGR_HIP(hipStreamSynchronize(ctx->stream));
    int32_t errv = 0;
    MGR_HIP(hipMemcpy(&errv, ctx->err_flag, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (errv) return trace_error(ctx, errv);
    std::vector<int32_t> ns(NQ);
    uint64_t sbits[2] = {0, 0};
    MGR_HIP(hipMemcpy(ns.data(), g.nseg, NQ * sizeof(int32_t), hipMemcpyDeviceToHost));
    MGR_HIP(hipMemcpy(sbits, ctx->shift_bits, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost));
    std::vector<int64_t> off(NQ + 1, 0);
    for (int64_t q = 0; q < NQ; q++) off[q + 1] = off[q] + ns[q];
    double smin, smax;
    std::memcpy(&smin, &sbits[0], sizeof(double));
    std::memcpy(&smax, &sbits[1], sizeof(double));
    g.shift_min = std::min(1.0, smin);
    g.shift_max = std::max(1.0, smax);
    if (off[NQ] > g.total_segments || !g.seg_cell) {
        for (void *p : {(void *)g.seg_cell, (void *)g.seg_shift, (void *)g.seg_dZ, (void *)g.seg_f32,
                        (void *)g.seg_hdz32})
            if (p) (void)hipFree(p);
        g.seg_cell = nullptr; g.seg_shift = nullptr; g.seg_dZ = nullptr; g.seg_f32 = nullptr; g.seg_hdz32 = nullptr;
        const int64_t T = std::max<int64_t>(off[NQ], 1);
        MGR_HIP(hipMalloc((void **)&g.seg_cell, T * sizeof(int32_t)));
        MGR_HIP(hipMalloc((void **)&g.seg_shift, T * sizeof(double)));
        MGR_HIP(hipMalloc((void **)&g.seg_dZ, T * sizeof(double)));
        MGR_HIP(hipMalloc((void **)&g.seg_f32, T * sizeof(int4)));
        MGR_HIP(hipMalloc((void **)&g.seg_hdz32, T * sizeof(float)));
        g.total_segments = T;
    }
    MGR_HIP(hipMemcpy(g.seg_off, off.data(), (NQ + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    auto fill = ctx->fwd_mask ? k_trace_rest<true, true> : k_trace_rest<true, false>;
    hipLaunchKernelGGL(fill, dim3(grid), dim3(128), 0, ctx->stream, N, (int)R, (int)ray, ctx->pos, ctx->vel,
                       ctx->nb_off, ctx->nb_rec, ctx->fwd_mask, ctx->is_boundary, ctx->rays,
                       ctx->d_antipod, ctx->dshift_max, nullptr, nullptr, g.seg_off, g.seg_cell, g.seg_shift,
                       g.seg_dZ, g.seg_f32, g.seg_hdz32, g.far_cell, g.far_shift, g.origin_shift, nullptr,
                       ctx->err_flag);
    MGR_LAUNCH_CHECK();
    MGR_HIP(hipStreamSynchronize(ctx->stream));
    g.ray = ray;
    return MGR_OK;
}

}  // namespace mgr

extern "C" {

int mgr_trace(mgr_ctx *ctx)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    MGR_HIP(hipSetDevice(ctx->device));
    return launch_trace(ctx);
}

int mgr_set_trace_options(mgr_ctx *ctx, int64_t record_steps, int candidate_masks)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    if (record_steps < -1 || record_steps > TRACE_RECORD_MAX || (record_steps > 0 && record_steps % 4))
        return ctx->fail(MGR_ERR_ARGUMENT, "record_steps: -1, 0 or a multiple of 4 up to 4096");
    MGR_HIP(hipSetDevice(ctx->device));
    ctx->trace_record_steps = record_steps < 0 ? 256 : record_steps;
    ctx->trace_masks = candidate_masks != 0;
    if (!ctx->trace_masks && ctx->fwd_mask) {
        MGR_HIP(hipStreamSynchronize(ctx->stream));
        (void)hipFree(ctx->fwd_mask);
        ctx->fwd_mask = nullptr;
    }
    return MGR_OK;
}

int mgr_get_ray(mgr_ctx *ctx, int64_t origin, int64_t ray, int64_t capacity, int64_t *cellNr,
                double *shift, double *dZ, int64_t *n_out)
{
    if (!ctx || !n_out || capacity < 0) return MGR_ERR_ARGUMENT;
    if (!ctx->traced) return ctx->fail(MGR_ERR_STATE, "mgr_trace first");
    if (origin < 0 || origin >= ctx->N) return ctx->fail(MGR_ERR_ARGUMENT, "origin out of range");
    {
        const int64_t k = std::upper_bound(ctx->m_cell_off.begin(), ctx->m_cell_off.end(), origin) -
                          ctx->m_cell_off.begin() - 1;
        if (ray < 0 || ray >= ctx->m_R[k]) return ctx->fail(MGR_ERR_ARGUMENT, "ray out of range");
    }
    if (capacity > 0 && (!cellNr || !shift || !dZ)) return MGR_ERR_ARGUMENT;
    MGR_HIP(hipSetDevice(ctx->device));
    const int64_t cap = std::max<int64_t>(capacity, 1);
    int64_t *dc = nullptr, *dn = nullptr;
    double *ds = nullptr, *dz = nullptr;
    MGR_HIP(hipMalloc((void **)&dc, cap * sizeof(int64_t)));
    MGR_HIP(hipMalloc((void **)&ds, cap * sizeof(double)));
    MGR_HIP(hipMalloc((void **)&dz, cap * sizeof(double)));
    MGR_HIP(hipMalloc((void **)&dn, sizeof(int64_t)));
    hipLaunchKernelGGL(k_expand_ray, dim3(1), dim3(64), 0, ctx->stream, (int32_t)origin,
                       ctx->h_cell_qbase[origin] + ray, ctx->seg_off, ctx->nseg, ctx->seg_cell, ctx->seg_shift,
                       ctx->seg_dZ, ctx->dshift_max, capacity, dc, ds, dz, dn);
    hipError_t le = hipGetLastError();
    hipError_t se = hipStreamSynchronize(ctx->stream);
    int rc = MGR_OK;
    if (le != hipSuccess || se != hipSuccess) rc = ctx->fail(MGR_ERR_DEVICE, "ray expansion kernel failed");
    if (rc == MGR_OK) {
        (void)hipMemcpy(n_out, dn, sizeof(int64_t), hipMemcpyDeviceToHost);
        const int64_t n = std::min<int64_t>(*n_out, capacity);
        if (n > 0) {
            (void)hipMemcpy(cellNr, dc, n * sizeof(int64_t), hipMemcpyDeviceToHost);
            (void)hipMemcpy(shift, ds, n * sizeof(double), hipMemcpyDeviceToHost);
            (void)hipMemcpy(dZ, dz, n * sizeof(double), hipMemcpyDeviceToHost);
        }
    }
    (void)hipFree(dc); (void)hipFree(ds); (void)hipFree(dz); (void)hipFree(dn);
    return rc;
}

int mgr_get_npoints_on_rays_rest(mgr_ctx *ctx, int64_t *npoints)
{
    if (!ctx || !npoints) return MGR_ERR_ARGUMENT;
    if (ctx->M != 1) return ctx->fail(MGR_ERR_UNSUPPORTED, "rest-frame traces are per model (batch of one)");
    MGR_HIP(hipSetDevice(ctx->device));
    const int mrc = mgr::launch_forward_masks(ctx);
    if (mrc) return mrc;
    const int64_t N = ctx->N, R = ctx->R, NQ = N * R;
    int32_t *d_ns = nullptr, *d_np = nullptr;
    MGR_HIP(hipMalloc((void **)&d_ns, std::max<int64_t>(NQ, 1) * sizeof(int32_t)));
    if (hipMalloc((void **)&d_np, std::max<int64_t>(NQ, 1) * sizeof(int32_t)) != hipSuccess) {
        (void)hipFree(d_ns);
        return ctx->fail(MGR_ERR_DEVICE, "out of device memory");
    }
    int rc = MGR_OK;
    std::vector<int32_t> np(NQ);
    int32_t errv = 0;
    if (hipMemsetAsync(ctx->err_flag, 0, sizeof(int32_t), ctx->stream) != hipSuccess) rc = MGR_ERR_DEVICE;
    if (!rc) {
        auto count = ctx->fwd_mask ? k_trace_rest<false, true> : k_trace_rest<false, false>;
        hipLaunchKernelGGL(count, dim3((unsigned)((NQ + 127) / 128)), dim3(128), 0, ctx->stream, N, (int)R, -1,
                           ctx->pos, ctx->vel, ctx->nb_off, ctx->nb_rec, ctx->fwd_mask,
                           ctx->is_boundary, ctx->rays, ctx->d_antipod, ctx->dshift_max, d_ns, d_np, nullptr, nullptr,
                           nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, ctx->err_flag);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess ||
            hipMemcpy(&errv, ctx->err_flag, sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(np.data(), d_np, NQ * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess)
            rc = MGR_ERR_DEVICE;
    }
    (void)hipFree(d_ns);
    (void)hipFree(d_np);
    if (rc) return ctx->fail(rc, "rest-frame ray count failed");
    if (errv) return trace_error(ctx, errv);
    for (int64_t o = 0; o < N; o++)
        for (int64_t r = 0; r < R / 2; r++)
            npoints[r * N + o] = (int64_t)np[o * R + r] + np[o * R + ctx->antipod[r]] + 1;
    return MGR_OK;
}

int mgr_get_npoints_on_rays(mgr_ctx *ctx, int64_t *npoints)
{
    if (!ctx || !npoints) return MGR_ERR_ARGUMENT;
    if (!ctx->traced) return ctx->fail(MGR_ERR_STATE, "mgr_trace first");
    MGR_HIP(hipSetDevice(ctx->device));
    const int64_t N = ctx->N, R = ctx->R;
    for (int64_t k = 0; k < ctx->M; k++)
        if (ctx->m_R[k] != R)
            return ctx->fail(MGR_ERR_UNSUPPORTED, "npoints_on_rays needs the same ray count in every model");
    std::vector<int32_t> np(ctx->NQ);
    MGR_HIP(hipMemcpy(np.data(), ctx->npts, ctx->NQ * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int64_t o = 0; o < N; o++) {
        const int64_t qb = ctx->h_cell_qbase[o];
        const int64_t k = std::upper_bound(ctx->m_cell_off.begin(), ctx->m_cell_off.end(), o) -
                          ctx->m_cell_off.begin() - 1;
        const int64_t rb = ctx->m_rbase[k];
        for (int64_t r = 0; r < R / 2; r++)
            npoints[r * N + o] = (int64_t)np[qb + r] + np[qb + ctx->antipod[rb + r] - rb] + 1;
    }
    return MGR_OK;
}

}  // extern "C"
