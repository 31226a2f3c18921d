// mgr_mlp.hip -- the learned warm start: a multilayer perceptron
// 8 -> 128 -> 128 -> 7 (GELU) that predicts per-cell log10 level-population
// fractions from local density, temperature, velocity gradient and a
// first-pass J, trained on populations the plain method converged.  It has no
// counterpart in the reference, whose starting point is LTE
// (Simulation::compute_LTE_level_populations, sim_lines.tpp:10-16); the plain
// LTE-started iteration stays the fallback, and the ALI loop that follows the
// warm start is unchanged, so converged populations are those of the plain
// method.
//
// Arithmetic: every matrix product runs on the CDNA4 matrix cores with
// bfloat16 operands, round to nearest even, and FP32 accumulation:
// v_mfma_f32_16x16x32_bf16 wherever K is a multiple of 32 (the hidden layer,
// the output layer, the back-propagation through W2 and the weight gradients,
// whose K is the tile's 64 rows), v_mfma_f32_16x16x16_bf16 for the two
// products with K = 16 (the 8 inputs and the 7 outputs, padded); biases,
// GELU, the loss and Adam run in FP32 on the FP32 master weights.
//
// Training runs without host round trips: minibatches are consecutive slices
// of a pseudo-random permutation of the rows computed on the device (a
// four-round Feistel network over the next even power of two with cycle
// walking, keyed per epoch), Adam's step count and the epoch's loss live in
// device memory, and each epoch is one hipGraph replay.
//
// Work mapping: a workgroup of 4 wavefronts takes tiles of 64 rows (cells);
// wave w owns rows 16w..16w+15 of the tile for the forward pass and the
// back-propagation of activations, keeping its pre-activations Z1, Z2 in
// registers in the MFMA result layout; the weight gradients, which reduce
// over rows, are split by output tile over the 4 waves with the 64 rows of
// the tile as the MFMA K dimension, and accumulate in registers over all the
// tiles a workgroup takes.  Per-workgroup partial gradients are reduced in a
// fixed order by the Adam kernel, so training is deterministic.
//
// MFMA operand layouts (lane l, r = l % 16, g = l / 16):
//   16x16x16: A[m = r][k = 4g + j], B[k = 4g + j][n = r], j = 0..3;
//   16x16x32: A[m = r][k = 8g + j], B[k = 8g + j][n = r], j = 0..7;
//   both:     D[m = 4g + j][n = r], j = 0..3.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mgr_common.hpp"

using namespace mgr;

namespace {

constexpr int IN = MGR_MLP_IN;        // 8 features
constexpr int HID = MGR_MLP_HIDDEN;   // 128
constexpr int OUT = MGR_MLP_OUT;      // 7 levels
constexpr int KIN = 16;               // IN padded to one MFMA K step
constexpr int NOUT = 16;              // OUT padded to one MFMA N tile
constexpr int ROWS = 64;              // rows per workgroup tile
constexpr int THREADS = 256;

// packed FP32 parameters: W1 [IN][HID], b1, W2 [HID][HID], b2, W3 [HID][OUT], b3
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + IN * HID;
constexpr int OFF_W2 = OFF_B1 + HID;
constexpr int OFF_B2 = OFF_W2 + HID * HID;
constexpr int OFF_W3 = OFF_B2 + HID;
constexpr int OFF_B3 = OFF_W3 + HID * OUT;
constexpr int NPAR = OFF_B3 + OUT;

// bf16 weight layouts, each laid out so that one lane reads its 4 operand
// values as one 8-byte word:
//   W1f [HID][KIN]   forward B of layer 1 (transposed, input padded)
//   W2f [HID][HID]   forward B of layer 2 (transposed)
//   W2b [HID][HID]   backward B of dH1 = dZ2 W2^T (as stored, [in][out])
//   W3f [NOUT][HID]  forward B of layer 3 (transposed, output padded)
//   W3b [HID][NOUT]  backward B of dH2 = dY W3^T ([in][out], output padded)
constexpr int BF_W1F = 0;
constexpr int BF_W2F = BF_W1F + HID * KIN;
constexpr int BF_W2B = BF_W2F + HID * HID;
constexpr int BF_W3F = BF_W2B + HID * HID;
constexpr int BF_W3B = BF_W3F + NOUT * HID;
constexpr int BF_TOTAL = BF_W3B + HID * NOUT;

// LDS of the training kernel (bf16 elements)
constexpr int L_X = BF_TOTAL;                 // [ROWS][KIN]
constexpr int L_H1 = L_X + ROWS * KIN;        // [ROWS][HID]
constexpr int L_H2 = L_H1 + ROWS * HID;       // [ROWS][HID]
constexpr int L_DY = L_H2 + ROWS * HID;       // [ROWS][NOUT]
constexpr int L_DZ2 = L_DY + ROWS * NOUT;     // [ROWS][HID]
constexpr int L_DZ1 = L_DZ2 + ROWS * HID;     // [ROWS][HID]
constexpr int L_TOTAL = L_DZ1 + ROWS * HID;
constexpr size_t LDS_TRAIN = (size_t)L_TOTAL * 2;   // 147,456 bytes

typedef short s4 __attribute__((ext_vector_type(4)));
typedef short s8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ s4 ld4(const uint16_t *p) { return *reinterpret_cast<const s4 *>(p); }
__device__ __forceinline__ s8 ld8(const uint16_t *p) { return *reinterpret_cast<const s8 *>(p); }
__device__ __forceinline__ s8 ld8s(const uint16_t *p, int stride)
{
    s8 r;
    for (int j = 0; j < 8; j++) r[j] = (short)p[j * stride];
    return r;
}
__device__ __forceinline__ f4 mfma(s4 a, s4 b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma32(s8 a, s8 b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a), __builtin_bit_cast(bf8, b), c, 0, 0, 0);
}

constexpr float SQRT_2_OVER_PI = 0.7978845608028654f;
constexpr float GELU_C = 0.044715f;

__device__ __forceinline__ float tanh_f(float u)
{
    // tanh u = 1 - 2 / (exp(2u) + 1), saturating for large |u|
    const float e = __expf(2.0f * fminf(fmaxf(u, -15.0f), 15.0f));
    return 1.0f - 2.0f / (e + 1.0f);
}
__device__ __forceinline__ float gelu(float z)
{
    const float t = tanh_f(SQRT_2_OVER_PI * (z + GELU_C * z * z * z));
    return 0.5f * z * (1.0f + t);
}
__device__ __forceinline__ float gelu_grad(float z)
{
    const float t = tanh_f(SQRT_2_OVER_PI * (z + GELU_C * z * z * z));
    const float du = SQRT_2_OVER_PI * (1.0f + 3.0f * GELU_C * z * z);
    return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * du;
}

struct Norm {
    float xm[IN], xs[IN], ym[OUT], ys[OUT];
};

// Device state of a training run: the epoch (counted from -1, advanced by
// k_epoch_begin at the start of every epoch replay), its permutation key
// and the shape of the permutation domain.
struct TrainState {
    int64_t  epoch;
    uint64_t seed, key;
    uint32_t n;              // rows permuted
    int      half;           // bits of each Feistel half
    uint32_t mask;           // (1 << half) - 1
};

// A four-round Feistel network on [0, 2^(2 half)): a bijection for any key.
__device__ __forceinline__ uint32_t feistel(uint32_t x, int half, uint32_t mask, uint64_t key)
{
    uint32_t L = x >> half, R = x & mask;
    for (int i = 0; i < 4; i++) {
        uint32_t f = R * 0x85EBCA6Bu + ((uint32_t)(key >> (16 * i)) ^ (0x9E3779B9u * (uint32_t)(i + 1)));
        f ^= f >> 13;
        f *= 0xC2B2AE35u;
        f ^= f >> 16;
        const uint32_t nl = R;
        R = (L ^ f) & mask;
        L = nl;
    }
    return (L << half) | R;
}

// Row at position i of the epoch's permutation of [0, n): the Feistel
// bijection, walked along its cycle until it lands inside [0, n) (the
// restriction of a permutation to a subset by cycle walking is a
// permutation of the subset).
__device__ __forceinline__ int64_t permuted_row(const TrainState &st, int64_t i)
{
    uint32_t x = (uint32_t)i;
    do { x = feistel(x, st.half, st.mask, st.key); } while (x >= st.n);
    return (int64_t)x;
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void k_epoch_begin(TrainState *st)
{
    st->epoch++;
    st->key = mix64(st->seed ^ mix64((uint64_t)st->epoch));
}

// Source row of batch row `row` (b0 + row of the epoch order, or idx[row],
// or row itself).
__device__ __forceinline__ int64_t source_row(const TrainState *st, const int64_t *idx, int64_t b0, int64_t row)
{
    if (st) return permuted_row(*st, b0 + row);
    return idx ? idx[row] : row;
}

// Copy the bf16 weight layouts (all of them, or the forward ones) into LDS.
__device__ __forceinline__ void load_weights(uint16_t *s, const uint16_t *__restrict__ wbf, bool fwd_only)
{
    const uint4 *src = reinterpret_cast<const uint4 *>(wbf);
    uint4 *dst = reinterpret_cast<uint4 *>(s);
    const int n = BF_TOTAL / 8;   // 8 bf16 per uint4
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int e = i * 8;
        if (fwd_only && e >= BF_W2B && e < BF_W3F) continue;
        if (fwd_only && e >= BF_W3B) continue;
        dst[i] = src[i];
    }
}

// Forward pass of the 16 rows of one wave (tile rows R0..R0+15).  sX holds
// the normalised, bf16 inputs.  Leaves H1, H2 in LDS, Z1, Z2 in registers
// and returns the output tile y (row R0 + 4g + j, column r).
__device__ __forceinline__ f4 forward_rows(const uint16_t *s, uint16_t *sH1, uint16_t *sH2, int R0,
                                           const float *__restrict__ P, f4 z1[8], f4 z2[8])
{
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const uint16_t *sX = s + L_X;
    const s4 ax = ld4(sX + (R0 + r) * KIN + 4 * g);
    for (int nt = 0; nt < 8; nt++) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma(ax, ld4(s + BF_W1F + (nt * 16 + r) * KIN + 4 * g), acc);
        const float b = P[OFF_B1 + nt * 16 + r];
        for (int j = 0; j < 4; j++) {
            acc[j] += b;
            sH1[(R0 + 4 * g + j) * HID + nt * 16 + r] = f2bf(gelu(acc[j]));
        }
        z1[nt] = acc;
    }
    __syncthreads();
    for (int nt = 0; nt < 8; nt++) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kt = 0; kt < 4; kt++)
            acc = mfma32(ld8(sH1 + (R0 + r) * HID + kt * 32 + 8 * g),
                         ld8(s + BF_W2F + (nt * 16 + r) * HID + kt * 32 + 8 * g), acc);
        const float b = P[OFF_B2 + nt * 16 + r];
        for (int j = 0; j < 4; j++) {
            acc[j] += b;
            sH2[(R0 + 4 * g + j) * HID + nt * 16 + r] = f2bf(gelu(acc[j]));
        }
        z2[nt] = acc;
    }
    __syncthreads();
    f4 y = {0.f, 0.f, 0.f, 0.f};
    for (int kt = 0; kt < 4; kt++)
        y = mfma32(ld8(sH2 + (R0 + r) * HID + kt * 32 + 8 * g), ld8(s + BF_W3F + r * HID + kt * 32 + 8 * g), y);
    const float b3 = r < OUT ? P[OFF_B3 + r] : 0.0f;
    for (int j = 0; j < 4; j++) y[j] += b3;
    return y;
}

// Stage the normalised inputs of tile rows into sX (zeros past n and in the
// padding columns).
__device__ __forceinline__ void stage_inputs(uint16_t *sX, const float *__restrict__ X, const TrainState *st,
                                             const int64_t *__restrict__ idx, int64_t b0, int64_t row0, int64_t n,
                                             const Norm &nm)
{
    for (int e = threadIdx.x; e < ROWS * KIN; e += blockDim.x) {
        const int rr = e / KIN, c = e - rr * KIN;
        const int64_t row = row0 + rr;
        float v = 0.0f;
        if (row < n && c < IN) {
            const int64_t src = source_row(st, idx, b0, row);
            v = (X[src * IN + c] - nm.xm[c]) / nm.xs[c];
        }
        sX[e] = f2bf(v);
    }
}

// Loss and gradients over rows [0, n) of a batch of X/T (rows b0 .. b0+n-1 of
// the epoch's permutation with st, else through idx when given, else the
// rows themselves): grid-stride over 64-row tiles; per-workgroup partial
// gradients (packed like the parameters) and partial loss go to
// part[blockIdx.x * (NPAR + 1)].
__global__ __launch_bounds__(THREADS) void k_mlp_grad(const float *__restrict__ X, const float *__restrict__ T,
                                                      const TrainState *st, const int64_t *__restrict__ idx,
                                                      int64_t b0, int64_t n, int64_t n_total,
                                                      const uint16_t *__restrict__ wbf, const float *__restrict__ P,
                                                      const Norm *__restrict__ norm, float *__restrict__ part)
{
    extern __shared__ __align__(16) uint16_t s[];
    uint16_t *sX = s + L_X, *sH1 = s + L_H1, *sH2 = s + L_H2, *sDY = s + L_DY, *sDZ2 = s + L_DZ2,
             *sDZ1 = s + L_DZ1;
    const Norm &nm = *norm;
    load_weights(s, wbf, false);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const int R0 = wave * 16;
    const float inv = 2.0f / ((float)n_total * (float)OUT);

    f4 gW2[16], gW1[2], gW3[2];
    for (int t = 0; t < 16; t++) gW2[t] = f4{0.f, 0.f, 0.f, 0.f};
    for (int t = 0; t < 2; t++) gW1[t] = gW3[t] = f4{0.f, 0.f, 0.f, 0.f};
    float gb = 0.0f, gb3 = 0.0f, loss = 0.0f;

    const int64_t ntiles = (n + ROWS - 1) / ROWS;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t row0 = tile * ROWS;
        __syncthreads();   // the previous tile's activations are consumed
        stage_inputs(sX, X, st, idx, b0, row0, n, nm);
        __syncthreads();
        f4 z1[8], z2[8];
        const f4 y = forward_rows(s, sH1, sH2, R0, P, z1, z2);
        // dY = 2 (y - t) / (n_total OUT) on valid rows and outputs
        for (int j = 0; j < 4; j++) {
            const int rr = R0 + 4 * g + j;
            const int64_t row = row0 + rr;
            float d = 0.0f;
            if (row < n && r < OUT) {
                const int64_t src = source_row(st, idx, b0, row);
                const float t = (T[src * OUT + r] - nm.ym[r]) / nm.ys[r];
                d = y[j] - t;
                loss += d * d;
                d *= inv;
            }
            sDY[rr * NOUT + r] = f2bf(d);
        }
        __syncthreads();
        // dH2 = dY W3^T, dZ2 = dH2 * gelu'(Z2)
        {
            const s4 a = ld4(sDY + (R0 + r) * NOUT + 4 * g);
            for (int nt = 0; nt < 8; nt++) {
                f4 acc = {0.f, 0.f, 0.f, 0.f};
                acc = mfma(a, ld4(s + BF_W3B + (nt * 16 + r) * NOUT + 4 * g), acc);
                for (int j = 0; j < 4; j++)
                    sDZ2[(R0 + 4 * g + j) * HID + nt * 16 + r] = f2bf(acc[j] * gelu_grad(z2[nt][j]));
            }
        }
        __syncthreads();
        // dH1 = dZ2 W2^T, dZ1 = dH1 * gelu'(Z1)
        for (int nt = 0; nt < 8; nt++) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
            for (int kt = 0; kt < 4; kt++)
                acc = mfma32(ld8(sDZ2 + (R0 + r) * HID + kt * 32 + 8 * g),
                             ld8(s + BF_W2B + (nt * 16 + r) * HID + kt * 32 + 8 * g), acc);
            for (int j = 0; j < 4; j++)
                sDZ1[(R0 + 4 * g + j) * HID + nt * 16 + r] = f2bf(acc[j] * gelu_grad(z1[nt][j]));
        }
        __syncthreads();
        // weight gradients over the 64 rows of the tile (MFMA K = rows)
        for (int ks = 0; ks < 2; ks++) {
            const int rb = 32 * ks + 8 * g;
            for (int q = 0; q < 2; q++) {
                const int it = 2 * wave + q;
                const s8 aH1 = ld8s(sH1 + rb * HID + it * 16 + r, HID);
                for (int ot = 0; ot < 8; ot++)
                    gW2[q * 8 + ot] = mfma32(aH1, ld8s(sDZ2 + rb * HID + ot * 16 + r, HID), gW2[q * 8 + ot]);
                gW1[q] = mfma32(ld8s(sX + rb * KIN + r, KIN), ld8s(sDZ1 + rb * HID + it * 16 + r, HID), gW1[q]);
                gW3[q] = mfma32(ld8s(sH2 + rb * HID + it * 16 + r, HID), ld8s(sDY + rb * NOUT + r, NOUT), gW3[q]);
            }
        }
        // bias gradients: column sums of dZ1 (threads 0..127), dZ2 (128..255), dY (0..6)
        {
            const int c = threadIdx.x & (HID - 1);
            const uint16_t *src = threadIdx.x < HID ? sDZ1 : sDZ2;
            for (int rr = 0; rr < ROWS; rr++) gb += bf2f(src[rr * HID + c]);
            if (threadIdx.x < OUT)
                for (int rr = 0; rr < ROWS; rr++) gb3 += bf2f(sDY[rr * NOUT + threadIdx.x]);
        }
    }

    // partial gradients of this workgroup, packed like the parameters
    float *out = part + (int64_t)blockIdx.x * (NPAR + 1);
    for (int q = 0; q < 2; q++) {
        const int it = 2 * wave + q;
        for (int ot = 0; ot < 8; ot++)
            for (int j = 0; j < 4; j++)
                out[OFF_W2 + (it * 16 + 4 * g + j) * HID + ot * 16 + r] = gW2[q * 8 + ot][j];
        for (int j = 0; j < 4; j++) {
            const int i = 4 * g + j;   // input feature (padded to 16)
            if (i < IN) out[OFF_W1 + i * HID + it * 16 + r] = gW1[q][j];
            if (r < OUT) out[OFF_W3 + (it * 16 + 4 * g + j) * OUT + r] = gW3[q][j];
        }
    }
    if (threadIdx.x < HID) out[OFF_B1 + threadIdx.x] = gb;
    else out[OFF_B2 + threadIdx.x - HID] = gb;
    if (threadIdx.x < OUT) out[OFF_B3 + threadIdx.x] = gb3;
    // loss: wave sums, then one lane per wave adds into LDS-free slots in order
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off, 64);
    __shared__ float wl[4];
    if (lane == 0) wl[wave] = loss;
    __syncthreads();
    if (threadIdx.x == 0) out[NPAR] = (wl[0] + wl[1]) + (wl[2] + wl[3]);
}

// bf16 copies of master parameter p in the weight layouts (the padding
// entries stay zero from the first k_mlp_pack).
__device__ __forceinline__ void repack_param(int p, float x, uint16_t *__restrict__ wbf)
{
    const uint16_t h = f2bf(x);
    if (p < OFF_B1) {                        // W1 [i][o]
        const int i = (p - OFF_W1) / HID, o = (p - OFF_W1) % HID;
        wbf[BF_W1F + o * KIN + i] = h;
    } else if (p >= OFF_W2 && p < OFF_B2) {  // W2 [i][o]
        const int k = p - OFF_W2, i = k / HID, o = k % HID;
        wbf[BF_W2F + o * HID + i] = h;
        wbf[BF_W2B + k] = h;
    } else if (p >= OFF_W3 && p < OFF_B3) {  // W3 [i][o]
        const int k = p - OFF_W3, i = k / OUT, o = k % OUT;
        wbf[BF_W3F + o * HID + i] = h;
        wbf[BF_W3B + i * NOUT + o] = h;
    }
}

// Sum the G partials of each parameter in a fixed order (8 interleaved
// sequential sums, then a fixed tree: deterministic), optionally apply one
// Adam step and refresh the parameter's bf16 copies.  Blocks of 256 threads
// cover 32 parameters (p = NPAR is the loss).  In a training run (st given)
// the step is step0 + epoch * nbatch + batch + 1 and the loss sum is added to
// loss_acc[epoch]; otherwise c1, c2 are given.
constexpr int RED_P = 32, RED_W = 8;
__global__ __launch_bounds__(RED_P * RED_W) void k_mlp_reduce_adam(
    const float *__restrict__ part, int G, float *__restrict__ grad_out, float *__restrict__ P,
    float *__restrict__ m, float *__restrict__ v, uint16_t *__restrict__ wbf, int do_adam, float lr, float c1,
    float c2, const TrainState *st, int64_t step0, int64_t nbatch, int64_t batch, float *__restrict__ loss_acc)
{
    __shared__ float sh[RED_W][RED_P];
    const int pi = threadIdx.x % RED_P, wi = threadIdx.x / RED_P;
    const int p = blockIdx.x * RED_P + pi;
    float s = 0.0f;
    if (p <= NPAR)
        for (int w = wi; w < G; w += RED_W) s += part[(int64_t)w * (NPAR + 1) + p];
    sh[wi][pi] = s;
    __syncthreads();
    if (wi != 0 || p > NPAR) return;
    const float gsum = ((sh[0][pi] + sh[1][pi]) + (sh[2][pi] + sh[3][pi])) +
                       ((sh[4][pi] + sh[5][pi]) + (sh[6][pi] + sh[7][pi]));
    if (grad_out) grad_out[p] = gsum;
    if (st && p == NPAR && loss_acc) loss_acc[st->epoch] += gsum;
    if (!do_adam || p == NPAR) return;
    if (st) {
        const float t = (float)(step0 + st->epoch * nbatch + batch + 1);
        c1 = 1.0f - powf(0.9f, t);
        c2 = 1.0f - powf(0.999f, t);
    }
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float mm = b1 * m[p] + (1.0f - b1) * gsum;
    const float vv = b2 * v[p] + (1.0f - b2) * gsum * gsum;
    m[p] = mm;
    v[p] = vv;
    const float x = P[p] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    P[p] = x;
    if (wbf) repack_param(p, x, wbf);
}

// FP32 master parameters -> the bf16 layouts.
__global__ void k_mlp_pack(const float *__restrict__ P, uint16_t *__restrict__ wbf)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= BF_TOTAL) return;
    float v = 0.0f;
    if (e < BF_W2F) {                       // W1f [HID][KIN]
        const int o = e / KIN, i = e % KIN;
        v = i < IN ? P[OFF_W1 + i * HID + o] : 0.0f;
    } else if (e < BF_W2B) {                // W2f [HID out][HID in]
        const int k = e - BF_W2F, o = k / HID, i = k % HID;
        v = P[OFF_W2 + i * HID + o];
    } else if (e < BF_W3F) {                // W2b [in][out]
        v = P[OFF_W2 + (e - BF_W2B)];
    } else if (e < BF_W3B) {                // W3f [NOUT][HID]
        const int k = e - BF_W3F, o = k / HID, i = k % HID;
        v = o < OUT ? P[OFF_W3 + i * OUT + o] : 0.0f;
    } else {                                // W3b [HID][NOUT]
        const int k = e - BF_W3B, i = k / NOUT, o = k % NOUT;
        v = o < OUT ? P[OFF_W3 + i * OUT + o] : 0.0f;
    }
    wbf[e] = f2bf(v);
}

// Forward only: Y (de-normalised) for rows [0, n), or, with pop != null, the
// level populations pop[p V + i] = pop_tot[p] 10^y_i / sum_i 10^y_i.
__global__ __launch_bounds__(THREADS) void k_mlp_forward(const float *__restrict__ X, int64_t n,
                                                         const uint16_t *__restrict__ wbf,
                                                         const float *__restrict__ P,
                                                         const Norm *__restrict__ norm, float *__restrict__ Y,
                                                         const double *__restrict__ pop_tot,
                                                         double *__restrict__ pop)
{
    extern __shared__ __align__(16) uint16_t s[];
    uint16_t *sX = s + L_X, *sH1 = s + L_H1, *sH2 = s + L_H2;
    const Norm &nm = *norm;
    load_weights(s, wbf, true);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const int R0 = wave * 16;
    __shared__ float sy[ROWS][NOUT];
    const int64_t ntiles = (n + ROWS - 1) / ROWS;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t row0 = tile * ROWS;
        __syncthreads();
        stage_inputs(sX, X, nullptr, nullptr, 0, row0, n, nm);
        __syncthreads();
        f4 z1[8], z2[8];
        const f4 y = forward_rows(s, sH1, sH2, R0, P, z1, z2);
        for (int j = 0; j < 4; j++) {
            const int rr = R0 + 4 * g + j;
            const float v = r < OUT ? y[j] * nm.ys[r] + nm.ym[r] : 0.0f;
            sy[rr][r] = v;
            const int64_t row = row0 + rr;
            if (Y && row < n && r < OUT) Y[row * OUT + r] = v;
        }
        __syncthreads();
        if (pop && threadIdx.x < ROWS) {
            const int64_t row = row0 + threadIdx.x;
            if (row < n) {
                double f[OUT], sum = 0.0;
                double mx = -1e300;
                for (int i = 0; i < OUT; i++) mx = fmax(mx, (double)sy[threadIdx.x][i]);
                for (int i = 0; i < OUT; i++) {
                    f[i] = exp10((double)sy[threadIdx.x][i] - mx);
                    sum += f[i];
                }
                for (int i = 0; i < OUT; i++) pop[row * OUT + i] = pop_tot[row] * (f[i] / sum);
            }
        }
    }
}

// The warm start's guard: a cell whose predicted populations are not all
// finite and positive, or give one of its lines an opacity <= 0 (the sign of
// n_j B_a - n_i B_s, as the emissivity/opacity kernel forms it), takes back
// the populations it held before the prediction.
__global__ void k_warm_guard(int64_t N, int V, int L, const int32_t *__restrict__ irad,
                             const int32_t *__restrict__ jrad, const double *__restrict__ Ba,
                             const double *__restrict__ Bs, const double *__restrict__ prev,
                             double *__restrict__ pop, unsigned long long *__restrict__ count)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    double *x = pop + p * V;
    bool ok = true;
    for (int i = 0; i < V; i++) ok = ok && x[i] > 0.0 && __builtin_isfinite(x[i]);
    for (int k = 0; k < L; k++) ok = ok && (x[jrad[k]] * Ba[k] - x[irad[k]] * Bs[k] > 0.0);
    if (!ok) {
        for (int i = 0; i < V; i++) x[i] = prev[p * V + i];
        atomicAdd(count, 1ull);
    }
}

// Per-cell features of the warm start (after a first radiation field and
// J_eff from the starting populations):
//   0 log10 n(collision partner 0)   1 log10 T   2 log10 n(species)
//   3 log10 max over neighbours |v_nb - v_p| / |x_nb - x_p|  (s^-1)
//   4..7 log10 (Jlin_k / B_nu_k(T)) for transitions k = 0..3 (0 if absent)
__global__ void k_features(int64_t N, int L, int nspecs, int num, int partner, const double *__restrict__ abundance,
                           const double *__restrict__ T, const double *__restrict__ pos,
                           const double *__restrict__ vel, const int64_t *__restrict__ nb_off,
                           const int32_t *__restrict__ nb, const double *__restrict__ frequency,
                           const double *__restrict__ Jlin, float *__restrict__ X)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const double floor_ = 1e-30;
    const double npart = abundance[p * nspecs + (partner >= 0 ? partner : num)];
    float *x = X + p * IN;
    x[0] = (float)log10(fmax(npart, floor_));
    x[1] = (float)log10(fmax(T[p], floor_));
    x[2] = (float)log10(fmax(abundance[p * nspecs + num], floor_));
    double gmax = 0.0;
    for (int64_t e = nb_off[p]; e < nb_off[p + 1]; e++) {
        const int64_t q = nb[e];
        double dv2 = 0.0, dx2 = 0.0;
        for (int d = 0; d < 3; d++) {
            const double dv = vel[3 * q + d] - vel[3 * p + d];
            const double dx = pos[3 * q + d] - pos[3 * p + d];
            dv2 += dv * dv;
            dx2 += dx * dx;
        }
        if (dx2 > 0.0) gmax = fmax(gmax, sqrt(dv2 / dx2));
    }
    x[3] = (float)log10(fmax(gmax * CC, floor_));
    for (int k = 0; k < 4; k++) {
        float f = 0.0f;
        if (k < L && Jlin) {   // Jlin null: no first pass, the inputs are 0
            const double nu = frequency[k];
            const double B = TWO_HH_OVER_CC_SQUARED * nu * nu * nu / expm1(HH_OVER_KB * nu / T[p]);
            f = (float)log10(fmax(Jlin[p * L + k] / B, floor_));
        }
        x[4 + k] = f;
    }
}

}  // namespace

struct mgr_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    float *P = nullptr, *m = nullptr, *v = nullptr;   // [NPAR] master and Adam moments
    uint16_t *wbf = nullptr;                          // [BF_TOTAL]
    Norm *norm = nullptr;
    Norm h_norm;
    int64_t step = 0;
    bool jlin_inputs = true;   // reads inputs 4..7 (Jlin / B_nu after a first pass)
    int fail(int code, const std::string &msg)
    {
        err = msg;
        return code;
    }
};

#define MLP_HIP(call)                                                                \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            mlp->fail(MGR_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
            return MGR_ERR_DEVICE;                                                   \
        }                                                                            \
    } while (0)

static int mlp_pack(mgr_mlp *mlp)
{
    hipLaunchKernelGGL(k_mlp_pack, dim3((BF_TOTAL + 255) / 256), dim3(256), 0, mlp->stream, mlp->P, mlp->wbf);
    MLP_HIP(hipGetLastError());
    return MGR_OK;
}

static int mlp_lds_attr(mgr_mlp *mlp)
{
    MLP_HIP(hipFuncSetAttribute((const void *)k_mlp_grad, hipFuncAttributeMaxDynamicSharedMemorySize, (int)LDS_TRAIN));
    MLP_HIP(hipFuncSetAttribute((const void *)k_mlp_forward, hipFuncAttributeMaxDynamicSharedMemorySize, (int)LDS_TRAIN));
    return MGR_OK;
}

extern "C" int64_t mgr_mlp_num_params(void) { return NPAR; }

extern "C" const char *mgr_mlp_last_error(const mgr_mlp *mlp) { return mlp ? mlp->err.c_str() : "null mlp"; }

extern "C" int mgr_mlp_create(int device, uint64_t seed, mgr_mlp **out)
{
    if (!out) return MGR_ERR_ARGUMENT;
    *out = nullptr;
    mgr_mlp *mlp = new mgr_mlp();
    mlp->device = device;
    auto bail = [&](int rc) {
        mgr_mlp_destroy(mlp);
        return rc;
    };
    if (hipSetDevice(device) != hipSuccess) return bail(MGR_ERR_DEVICE);
    if (hipStreamCreate(&mlp->stream) != hipSuccess) return bail(MGR_ERR_DEVICE);
    if (hipMalloc(&mlp->P, NPAR * sizeof(float)) != hipSuccess || hipMalloc(&mlp->m, NPAR * sizeof(float)) != hipSuccess ||
        hipMalloc(&mlp->v, NPAR * sizeof(float)) != hipSuccess ||
        hipMalloc(&mlp->wbf, BF_TOTAL * sizeof(uint16_t)) != hipSuccess ||
        hipMalloc(&mlp->norm, sizeof(Norm)) != hipSuccess)
        return bail(MGR_ERR_NOMEM);
    // He-style initialisation from a 64-bit LCG + Box-Muller, zero biases
    std::vector<float> P(NPAR, 0.0f);
    uint64_t st = seed * 6364136223846793005ULL + 1442695040888963407ULL;
    auto uni = [&]() {
        st = st * 6364136223846793005ULL + 1442695040888963407ULL;
        return ((st >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    auto gauss = [&]() { return std::sqrt(-2.0 * std::log(uni())) * std::cos(2.0 * PI * uni()); };
    for (int i = 0; i < IN * HID; i++) P[OFF_W1 + i] = (float)(gauss() * std::sqrt(2.0 / IN));
    for (int i = 0; i < HID * HID; i++) P[OFF_W2 + i] = (float)(gauss() * std::sqrt(2.0 / HID));
    for (int i = 0; i < HID * OUT; i++) P[OFF_W3 + i] = (float)(gauss() * std::sqrt(1.0 / HID));
    for (int i = 0; i < IN; i++) { mlp->h_norm.xm[i] = 0.0f; mlp->h_norm.xs[i] = 1.0f; }
    for (int i = 0; i < OUT; i++) { mlp->h_norm.ym[i] = 0.0f; mlp->h_norm.ys[i] = 1.0f; }
    if (hipMemcpy(mlp->P, P.data(), NPAR * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(mlp->m, 0, NPAR * sizeof(float)) != hipSuccess ||
        hipMemset(mlp->v, 0, NPAR * sizeof(float)) != hipSuccess ||
        hipMemcpy(mlp->norm, &mlp->h_norm, sizeof(Norm), hipMemcpyHostToDevice) != hipSuccess)
        return bail(MGR_ERR_DEVICE);
    int rc = mlp_lds_attr(mlp);
    if (!rc) rc = mlp_pack(mlp);
    if (!rc && hipStreamSynchronize(mlp->stream) != hipSuccess) rc = MGR_ERR_DEVICE;
    if (rc) return bail(rc);
    *out = mlp;
    return MGR_OK;
}

extern "C" void mgr_mlp_destroy(mgr_mlp *mlp)
{
    if (!mlp) return;
    (void)hipSetDevice(mlp->device);
    if (mlp->stream) (void)hipStreamSynchronize(mlp->stream);
    if (mlp->P) (void)hipFree(mlp->P);
    if (mlp->m) (void)hipFree(mlp->m);
    if (mlp->v) (void)hipFree(mlp->v);
    if (mlp->wbf) (void)hipFree(mlp->wbf);
    if (mlp->norm) (void)hipFree(mlp->norm);
    if (mlp->stream) (void)hipStreamDestroy(mlp->stream);
    delete mlp;
}

extern "C" int mgr_mlp_set_params(mgr_mlp *mlp, const float *params)
{
    if (!mlp || !params) return MGR_ERR_ARGUMENT;
    MLP_HIP(hipSetDevice(mlp->device));
    MLP_HIP(hipMemcpy(mlp->P, params, NPAR * sizeof(float), hipMemcpyHostToDevice));
    MLP_HIP(hipMemset(mlp->m, 0, NPAR * sizeof(float)));
    MLP_HIP(hipMemset(mlp->v, 0, NPAR * sizeof(float)));
    mlp->step = 0;
    int rc = mlp_pack(mlp);
    if (rc) return rc;
    MLP_HIP(hipStreamSynchronize(mlp->stream));
    return MGR_OK;
}

extern "C" int mgr_mlp_get_params(mgr_mlp *mlp, float *params)
{
    if (!mlp || !params) return MGR_ERR_ARGUMENT;
    MLP_HIP(hipSetDevice(mlp->device));
    MLP_HIP(hipStreamSynchronize(mlp->stream));
    MLP_HIP(hipMemcpy(params, mlp->P, NPAR * sizeof(float), hipMemcpyDeviceToHost));
    return MGR_OK;
}

extern "C" int mgr_mlp_set_normalisation(mgr_mlp *mlp, const float *x_mean, const float *x_std, const float *y_mean,
                                         const float *y_std)
{
    if (!mlp || !x_mean || !x_std || !y_mean || !y_std) return MGR_ERR_ARGUMENT;
    for (int i = 0; i < IN; i++)
        if (!(x_std[i] > 0.0f)) return mlp->fail(MGR_ERR_ARGUMENT, "x_std must be positive");
    for (int i = 0; i < OUT; i++)
        if (!(y_std[i] > 0.0f)) return mlp->fail(MGR_ERR_ARGUMENT, "y_std must be positive");
    std::memcpy(mlp->h_norm.xm, x_mean, sizeof(mlp->h_norm.xm));
    std::memcpy(mlp->h_norm.xs, x_std, sizeof(mlp->h_norm.xs));
    std::memcpy(mlp->h_norm.ym, y_mean, sizeof(mlp->h_norm.ym));
    std::memcpy(mlp->h_norm.ys, y_std, sizeof(mlp->h_norm.ys));
    MLP_HIP(hipSetDevice(mlp->device));
    MLP_HIP(hipMemcpy(mlp->norm, &mlp->h_norm, sizeof(Norm), hipMemcpyHostToDevice));
    return MGR_OK;
}

static int grid_for_rows(int64_t n)
{
    const int64_t tiles = (n + ROWS - 1) / ROWS;
    return (int)(tiles < 1024 ? (tiles > 0 ? tiles : 1) : 1024);
}

extern "C" int mgr_mlp_forward(mgr_mlp *mlp, const float *X, int64_t n, float *Y)
{
    if (!mlp || !X || !Y || n < 0) return MGR_ERR_ARGUMENT;
    if (n == 0) return MGR_OK;
    MLP_HIP(hipSetDevice(mlp->device));
    float *dX = nullptr, *dY = nullptr;
    MLP_HIP(hipMalloc(&dX, n * IN * sizeof(float)));
    if (hipMalloc(&dY, n * OUT * sizeof(float)) != hipSuccess) {
        (void)hipFree(dX);
        return mlp->fail(MGR_ERR_NOMEM, "forward: output allocation");
    }
    int rc = MGR_OK;
    if (hipMemcpy(dX, X, n * IN * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) rc = MGR_ERR_DEVICE;
    if (!rc) {
        hipLaunchKernelGGL(k_mlp_forward, dim3(grid_for_rows(n)), dim3(THREADS), LDS_TRAIN, mlp->stream, dX, n,
                           mlp->wbf, mlp->P, mlp->norm, dY, (const double *)nullptr, (double *)nullptr);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(mlp->stream) != hipSuccess) rc = MGR_ERR_DEVICE;
    }
    if (!rc && hipMemcpy(Y, dY, n * OUT * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) rc = MGR_ERR_DEVICE;
    (void)hipFree(dX);
    (void)hipFree(dY);
    if (rc) return mlp->fail(rc, "forward failed");
    return MGR_OK;
}

// one gradient evaluation of rows [0, n) of device-resident X/T (no Adam step)
static int mlp_grad_device(mgr_mlp *mlp, const float *dX, const float *dT, int64_t n, float *dpart, int G,
                           float *dgrad)
{
    hipLaunchKernelGGL(k_mlp_grad, dim3(G), dim3(THREADS), LDS_TRAIN, mlp->stream, dX, dT,
                       (const TrainState *)nullptr, (const int64_t *)nullptr, (int64_t)0, n, n, mlp->wbf, mlp->P,
                       mlp->norm, dpart);
    MLP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((NPAR + 1 + RED_P - 1) / RED_P), dim3(RED_P * RED_W), 0, mlp->stream,
                       dpart, G, dgrad, mlp->P, mlp->m, mlp->v, (uint16_t *)nullptr, 0, 0.0f, 1.0f, 1.0f,
                       (const TrainState *)nullptr, (int64_t)0, (int64_t)0, (int64_t)0, (float *)nullptr);
    MLP_HIP(hipGetLastError());
    return MGR_OK;
}

extern "C" int mgr_mlp_loss_grad(mgr_mlp *mlp, const float *X, const float *T, int64_t n, float *loss, float *grad)
{
    if (!mlp || !X || !T || n <= 0) return MGR_ERR_ARGUMENT;
    MLP_HIP(hipSetDevice(mlp->device));
    const int G = grid_for_rows(n);
    float *dX = nullptr, *dT = nullptr, *dpart = nullptr, *dgrad = nullptr;
    int rc = MGR_OK;
    if (hipMalloc(&dX, n * IN * sizeof(float)) != hipSuccess || hipMalloc(&dT, n * OUT * sizeof(float)) != hipSuccess ||
        hipMalloc(&dpart, (size_t)G * (NPAR + 1) * sizeof(float)) != hipSuccess ||
        hipMalloc(&dgrad, (NPAR + 1) * sizeof(float)) != hipSuccess)
        rc = MGR_ERR_NOMEM;
    if (!rc && (hipMemcpy(dX, X, n * IN * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(dT, T, n * OUT * sizeof(float), hipMemcpyHostToDevice) != hipSuccess))
        rc = MGR_ERR_DEVICE;
    if (!rc) rc = mlp_grad_device(mlp, dX, dT, n, dpart, G, dgrad);
    std::vector<float> h(NPAR + 1);
    if (!rc && (hipStreamSynchronize(mlp->stream) != hipSuccess ||
                hipMemcpy(h.data(), dgrad, (NPAR + 1) * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess))
        rc = MGR_ERR_DEVICE;
    (void)hipFree(dX); (void)hipFree(dT); (void)hipFree(dpart); (void)hipFree(dgrad);
    if (rc) return mlp->fail(rc, "loss_grad failed");
    if (grad) std::memcpy(grad, h.data(), NPAR * sizeof(float));
    if (loss) *loss = h[NPAR] / (float)(n * OUT);
    return MGR_OK;
}

// One epoch = minibatches b0 = 0, batch, 2 batch, ... of the epoch's row
// permutation, each a gradient and an Adam step with the bf16 repack; captured
// once as a hipGraph and replayed per epoch, with no host synchronisation
// until the end.
extern "C" int mgr_mlp_train(mgr_mlp *mlp, const float *X, const float *T, int64_t n, int64_t epochs, int64_t batch,
                             float lr, uint64_t seed, float *loss_per_epoch)
{
    if (!mlp || !X || !T || n <= 0 || epochs <= 0 || batch <= 0 || !(lr > 0.0f)) return MGR_ERR_ARGUMENT;
    if (n > (int64_t)1 << 31) return mlp->fail(MGR_ERR_ARGUMENT, "more than 2^31 training rows");
    MLP_HIP(hipSetDevice(mlp->device));
    const int64_t nbatch = (n + batch - 1) / batch;
    const int Gmax = grid_for_rows(std::min(batch, n));
    float *dX = nullptr, *dT = nullptr, *dpart = nullptr, *dloss = nullptr;
    TrainState *dst = nullptr;
    int rc = MGR_OK;
    if (hipMalloc(&dX, n * IN * sizeof(float)) != hipSuccess || hipMalloc(&dT, n * OUT * sizeof(float)) != hipSuccess ||
        hipMalloc(&dpart, (size_t)Gmax * (NPAR + 1) * sizeof(float)) != hipSuccess ||
        hipMalloc(&dloss, epochs * sizeof(float)) != hipSuccess || hipMalloc(&dst, sizeof(TrainState)) != hipSuccess)
        rc = MGR_ERR_NOMEM;
    TrainState h{};
    h.epoch = -1;
    h.seed = seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    h.n = (uint32_t)n;
    int bits = 2;
    while (((int64_t)1 << bits) < n) bits += 2;
    h.half = bits / 2;
    h.mask = (1u << h.half) - 1u;
    if (!rc && (hipMemcpy(dX, X, n * IN * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(dT, T, n * OUT * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(dst, &h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemset(dloss, 0, epochs * sizeof(float)) != hipSuccess))
        rc = MGR_ERR_DEVICE;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (!rc && hipStreamBeginCapture(mlp->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) rc = MGR_ERR_DEVICE;
    if (!rc) {
        hipLaunchKernelGGL(k_epoch_begin, dim3(1), dim3(1), 0, mlp->stream, dst);
        for (int64_t b = 0; b < nbatch; b++) {
            const int64_t b0 = b * batch, nb = std::min(batch, n - b0);
            const int G = grid_for_rows(nb);
            hipLaunchKernelGGL(k_mlp_grad, dim3(G), dim3(THREADS), LDS_TRAIN, mlp->stream, dX, dT,
                               (const TrainState *)dst, (const int64_t *)nullptr, b0, nb, nb, mlp->wbf, mlp->P,
                               mlp->norm, dpart);
            hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((NPAR + 1 + RED_P - 1) / RED_P), dim3(RED_P * RED_W), 0,
                               mlp->stream, dpart, G, (float *)nullptr, mlp->P, mlp->m, mlp->v, mlp->wbf, 1, lr, 1.0f,
                               1.0f, (const TrainState *)dst, mlp->step, nbatch, b, dloss);
        }
        const hipError_t le = hipGetLastError();
        const hipError_t ce = hipStreamEndCapture(mlp->stream, &graph);
        if (le != hipSuccess || ce != hipSuccess ||
            hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
            rc = MGR_ERR_DEVICE;
    }
    for (int64_t e = 0; !rc && e < epochs; e++)
        if (hipGraphLaunch(exec, mlp->stream) != hipSuccess) rc = MGR_ERR_DEVICE;
    if (!rc && hipStreamSynchronize(mlp->stream) != hipSuccess) rc = MGR_ERR_DEVICE;
    if (!rc) mlp->step += epochs * nbatch;
    if (!rc && loss_per_epoch) {
        std::vector<float> l(epochs);
        if (hipMemcpy(l.data(), dloss, epochs * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) rc = MGR_ERR_DEVICE;
        for (int64_t e = 0; !rc && e < epochs; e++) loss_per_epoch[e] = (float)((double)l[e] / ((double)n * OUT));
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipFree(dX); (void)hipFree(dT); (void)hipFree(dpart); (void)hipFree(dloss); (void)hipFree(dst);
    if (rc) return mlp->fail(rc, "train failed");
    return MGR_OK;
}

// ---- context side: features and warm start -----------------------------------

namespace mgr {
int launch_features(mgr_ctx *ctx, float *dX, bool jlin);
}

int mgr::launch_features(mgr_ctx *ctx, float *dX, bool jlin)
{
    const int partner = ctx->h_cp_num.empty() ? -1 : (int)ctx->h_cp_num[0];
    hipLaunchKernelGGL(k_features, dim3((unsigned)((ctx->N + 255) / 256)), dim3(256), 0, ctx->stream, ctx->N,
                       (int)ctx->L, (int)ctx->nspecs, (int)ctx->num, partner, ctx->abundance, ctx->temperature,
                       ctx->pos, ctx->vel, ctx->nb_off, ctx->nb, ctx->frequency, jlin ? ctx->Jlin : nullptr, dX);
    MGR_LAUNCH_CHECK();
    return MGR_OK;
}

extern "C" int mgr_features(mgr_ctx *ctx, float *X)
{
    if (!ctx || !X) return MGR_ERR_ARGUMENT;
    if (ctx->S > 1) return ctx->fail(MGR_ERR_UNSUPPORTED, "the warm-start features are defined for one line-producing species");
    MGR_HIP(hipSetDevice(ctx->device));
    float *dX = nullptr;
    MGR_HIP(hipMalloc(&dX, ctx->N * IN * sizeof(float)));
    int rc = mgr::launch_features(ctx, dX, ctx->J_set);   // before any radiation field the J inputs are 0
    if (!rc && (hipStreamSynchronize(ctx->stream) != hipSuccess ||
                hipMemcpy(X, dX, ctx->N * IN * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess))
        rc = ctx->fail(MGR_ERR_DEVICE, "features copy failed");
    (void)hipFree(dX);
    return rc;
}

extern "C" int mgr_warm_start(mgr_ctx *ctx, mgr_mlp *mlp)
{
    if (!ctx || !mlp) return MGR_ERR_ARGUMENT;
    if (ctx->V != OUT || ctx->S > 1)
        return ctx->fail(MGR_ERR_UNSUPPORTED, "the warm-start network predicts the 7 levels of one species");
    if (mlp->jlin_inputs && !ctx->J_set)
        return ctx->fail(MGR_ERR_STATE, "warm start needs a radiation field and J_eff first (the network reads J)");
    if (ctx->device != mlp->device) return ctx->fail(MGR_ERR_ARGUMENT, "context and network on different devices");
    MGR_HIP(hipSetDevice(ctx->device));
    MGR_HIP(hipStreamSynchronize(mlp->stream));
    ctx->warm_fallbacks = 0;
    float *dX = nullptr;
    double *dPrev = nullptr;
    unsigned long long *dCount = nullptr;
    const size_t npop = (size_t)ctx->N * OUT;
    if (hipMalloc(&dX, ctx->N * IN * sizeof(float)) != hipSuccess ||
        hipMalloc(&dPrev, npop * sizeof(double)) != hipSuccess ||
        hipMalloc(&dCount, sizeof(unsigned long long)) != hipSuccess) {
        (void)hipFree(dX);
        (void)hipFree(dPrev);
        return ctx->fail(MGR_ERR_NOMEM, "warm start: device allocation failed");
    }
    int rc = mgr::launch_features(ctx, dX, mlp->jlin_inputs);
    if (!rc && (hipMemcpyAsync(dPrev, ctx->pop, npop * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess ||
                hipMemsetAsync(dCount, 0, sizeof(unsigned long long), ctx->stream) != hipSuccess))
        rc = ctx->fail(MGR_ERR_DEVICE, "warm start: copy failed");
    if (!rc) {
        hipLaunchKernelGGL(k_mlp_forward, dim3(grid_for_rows(ctx->N)), dim3(THREADS), LDS_TRAIN, ctx->stream, dX,
                           ctx->N, mlp->wbf, mlp->P, mlp->norm, (float *)nullptr, ctx->pop_tot, ctx->pop);
        hipLaunchKernelGGL(k_warm_guard, dim3((unsigned)((ctx->N + 255) / 256)), dim3(256), 0, ctx->stream, ctx->N,
                           (int)OUT, (int)ctx->L, ctx->irad, ctx->jrad, ctx->Ba, ctx->Bs, dPrev, ctx->pop, dCount);
        if (hipGetLastError() != hipSuccess) rc = ctx->fail(MGR_ERR_DEVICE, "warm-start launch failed");
    }
    if (!rc) rc = launch_emissivity_opacity(ctx);
    unsigned long long nfb = 0;
    if (!rc && (hipMemcpyAsync(&nfb, dCount, sizeof(nfb), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
                hipStreamSynchronize(ctx->stream) != hipSuccess))
        rc = ctx->fail(MGR_ERR_DEVICE, "warm start failed");
    (void)hipFree(dX);
    (void)hipFree(dPrev);
    (void)hipFree(dCount);
    if (!rc) {
        ctx->warm_fallbacks = (int64_t)nfb;
        ctx->pops_set = true;
    }
    return rc;
}

extern "C" int mgr_mlp_set_jlin_inputs(mgr_mlp *mlp, int use)
{
    if (!mlp) return MGR_ERR_ARGUMENT;
    mlp->jlin_inputs = use != 0;
    return MGR_OK;
}

extern "C" int mgr_warm_start_fallbacks(const mgr_ctx *ctx, int64_t *cells)
{
    if (!ctx || !cells) return MGR_ERR_ARGUMENT;
    *cells = ctx->warm_fallbacks;
    return MGR_OK;
}
