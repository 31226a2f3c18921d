// mgr_levels.hip -- level populations: J_eff, the per-cell statistical
// equilibrium (FP64 LU), convergence statistics, Ng acceleration and the ALI
// driver that replays one hipGraph per iteration.
//
// Replaces Simulation::compute_Jeff (src/Simulation/sim_lines.tpp:176-220),
// LineProducingSpecies::update_using_statistical_equilibrium
// (lineProducingSpecies.tpp:249-495), ::check_for_convergence (:120-171),
// ::update_using_Ng_acceleration (:181-238) and
// Simulation::compute_level_populations (sim_lines.tpp:61-164).
//
// With n_off_diag = 0 every rate-equation triplet couples levels of a single
// cell, so the reference's global SparseLU<COLAMD> of size (N V)^2 is block
// diagonal: here one thread solves one cell's V x V system with partial
// pivoting in FP64 registers (V known at compile time for V <= 8).
#include <cmath>
#include <mutex>
#include <string>

#include "mgr_common.hpp"

using namespace mgr;

namespace {

constexpr int32_t ERR_SINGULAR = 16;
constexpr int32_t ERR_NONFINITE = 64;   // a level population is NaN or infinite

// The Lambda* diagonal of RayPair::update_Lambda (raypair_lambda.tpp:94-149),
// n_off_diag = 0: Ls(p,k) = sum over bins f whose back-map names transition k
// of A_k w_z nu phi_k(nu) * [sum_r 2 w_r L_diag / chi](p,f).  The
// radiation-field kernel leaves each 64-bin wave's share of every sum in
// Lpart; here they are added in chunk order.
__global__ void k_lambda_sum(int64_t N, int L, int nchunk, const double *__restrict__ Lpart,
                             double *__restrict__ Ls)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * L) return;
    const int64_t p = t / L;
    const int k = (int)(t - p * L);
    double ls = 0.0;
    for (int c = 0; c < nchunk; c++) ls += Lpart[(p * nchunk + c) * L + k];
    Ls[t] = ls;
}

// Simulation::compute_Jeff (sim_lines.tpp:176-220): Jlin = sum_z w_z J(p, nr_line),
// Jeff = Jlin - h/(4 pi) Ls n_i, Jdif = h/(4 pi) Ls n_i.
// With band elements (band_cell non-null) every element of (p, k) enters:
// diff = Ls(p,k) n_i(p) + sum over the slots of p of Ls(slot, k) n_i(cell).
// One species: its L lines are lines loff .. loff+L-1 of the per-line arrays
// [N][Lst]; pop, qw and irad are the species' own.
__global__ void k_jeff(int64_t N, int L, int Lst, int loff, int Q, int F, int V, const double *__restrict__ J,
                       const int32_t *__restrict__ nr_line, const double *__restrict__ qw,
                       const int32_t *__restrict__ irad, const double *__restrict__ pop,
                       const double *__restrict__ Ls, double *__restrict__ Jlin,
                       double *__restrict__ Jeff, double *__restrict__ Jdif, int64_t S,
                       const int32_t *__restrict__ band_cell, const double *__restrict__ band_Ls)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * L) return;
    const int64_t p = t / L;
    const int k = (int)(t % L);
    const int64_t pk = p * Lst + loff + k;
    double jl = 0.0;
    for (int z = 0; z < Q; z++) jl += qw[z] * J[p * F + nr_line[pk * Q + z]];
    double diff = Ls[pk] * pop[p * V + irad[k]];
    if (band_cell)
        for (int64_t sl = p * S; sl < (p + 1) * S; sl++) {
            const int32_t c = band_cell[sl];
            if (c >= 0) diff += band_Ls[sl * Lst + loff + k] * pop[(int64_t)c * V + irad[k]];
        }
    Jlin[pk] = jl;
    Jeff[pk] = jl - HH_OVER_FOUR_PI * diff;
    Jdif[pk] = HH_OVER_FOUR_PI * diff;
}

// search (src/Functions/interpolation.tpp:14-54)
__device__ __forceinline__ int search_table(const double *x, int n, double value)
{
    int start = 0, stop = n - 1;
    if (value >= x[stop]) return stop;
    else if (value <= x[start]) return start;
    while (stop > start + 1) {
        const int middle = (stop + start) / 2;
        if (value > x[middle]) start = middle;
        else if (value < x[middle]) stop = middle;
        else return middle;
    }
    return stop;
}

// One species' rate equations: its V levels, L lines (line k of its own at
// p*Lst + k of Jeff and Ls) and ncolpar collision partners; the arrays point
// at the species' own parts.
struct StateqArgs {
    int64_t N;
    int V, L, Lst, nspecs, ncolpar;
    const double *Jeff, *Ls, *A, *Bs, *Ba, *abundance, *T, *pop_tot;
    const int32_t *irad, *jrad;
    const int32_t *cp_num, *cp_ncol, *cp_ntmp, *cp_koff, *cp_toff, *cp_coff, *cp_op, *icol, *jcol;
    const double *tmp, *Cd, *Ce;
    double *pop, *prev1, *prev2, *prev3;
    int32_t *err_flag;
    const int32_t *cell_model, *m_active;   // converged models of a batch keep their populations
};

// History shift and the cell's own triplets (lineProducingSpecies.tpp:260-388)
// in the transpose convention of the reference, last row = conservation.
template <int VV>
__device__ __forceinline__ void stateq_assemble(const StateqArgs &a, int64_t p, int V, double *M, double *y)
{
    const int Vn = (VV > 0) ? VV : V;
    // history: prev3 <- prev2 <- prev1 <- population (lineProducingSpecies.tpp:260-262)
    for (int i = 0; i < Vn; i++) {
        const int64_t ind = p * Vn + i;
        const double x = a.pop[ind];
        a.prev3[ind] = a.prev2[ind];
        a.prev2[ind] = a.prev1[ind];
        a.prev1[ind] = x;
        y[i] = 0.0;
        for (int j = 0; j < Vn; j++) M[i * Vn + j] = 0.0;
    }
    const double *pop = a.prev1 + p * Vn;
    // radiative transitions (transpose convention of the reference)
    for (int k = 0; k < a.L; k++) {
        const double Je = a.Jeff[p * a.Lst + k];
        const double v_IJ = a.A[k] + a.Bs[k] * Je;
        const double v_JI = a.Ba[k] * Je;
        const int I = a.irad[k], Jl = a.jrad[k];
        if (Jl != Vn - 1) { M[Jl * Vn + I] += v_IJ; M[Jl * Vn + Jl] += -v_JI; }
        if (I != Vn - 1) { M[I * Vn + Jl] += v_JI; M[I * Vn + I] += -v_IJ; }
    }
    // approximated Lambda operator with the pre-solve opacity
    for (int k = 0; k < a.L; k++) {
        const int I = a.irad[k], Jl = a.jrad[k];
        const double opacity = HH_OVER_FOUR_PI * (pop[Jl] * a.Ba[k] - pop[I] * a.Bs[k]);
        const double v_IJ = -a.Ls[p * a.Lst + k] * opacity;
        if (Jl != Vn - 1) M[Jl * Vn + I] += v_IJ;
        if (I != Vn - 1) M[I * Vn + I] += -v_IJ;
    }
    // collisional transitions (collisionPartner.tpp:15-72)
    const double Tp = a.T[p];
    for (int c = 0; c < a.ncolpar; c++) {
        double abn = a.abundance[p * a.nspecs + a.cp_num[c]];
        if (a.cp_op[c] != 0) {
            const double frac_H2_para = 1.0 / (1.0 + 9.0 * exp(-170.5 / Tp));
            if (a.cp_op[c] == 1) abn *= (1.0 - frac_H2_para);
            if (a.cp_op[c] == 2) abn *= frac_H2_para;
        }
        const int ncol = a.cp_ncol[c], ntmp = a.cp_ntmp[c];
        const double *tmp = a.tmp + a.cp_toff[c];
        const double *Cd = a.Cd + a.cp_coff[c];
        const double *Ce = a.Ce + a.cp_coff[c];
        const int t = search_table(tmp, ntmp, Tp);
        const double step = (t == 0) ? 0.0 : (Tp - tmp[t - 1]) / (tmp[t] - tmp[t - 1]);
        for (int k = 0; k < ncol; k++) {
            double ce, cd;
            if (t == 0) {
                ce = Ce[k];
                cd = Cd[k];
            } else {
                ce = Ce[(t - 1) * ncol + k] + (Ce[t * ncol + k] - Ce[(t - 1) * ncol + k]) * step;
                cd = Cd[(t - 1) * ncol + k] + (Cd[t * ncol + k] - Cd[(t - 1) * ncol + k]) * step;
            }
            const double v_IJ = cd * abn, v_JI = ce * abn;
            const int I = a.icol[a.cp_koff[c] + k], Jl = a.jcol[a.cp_koff[c] + k];
            if (Jl != Vn - 1) { M[Jl * Vn + I] += v_IJ; M[Jl * Vn + Jl] += -v_JI; }
            if (I != Vn - 1) { M[I * Vn + Jl] += v_JI; M[I * Vn + I] += -v_IJ; }
        }
    }
    // conservation replaces the last level's equation
    for (int i = 0; i < Vn; i++) M[(Vn - 1) * Vn + i] += 1.0;
    y[Vn - 1] = a.pop_tot[p];
}

template <int VV>
__device__ __forceinline__ void stateq_cell(const StateqArgs &a, int64_t p, int V, double *M, double *y)
{
    const int Vn = (VV > 0) ? VV : V;
    stateq_assemble<VV>(a, p, Vn, M, y);
    // LU with partial pivoting
    for (int c = 0; c < Vn; c++) {
        int piv = c;
        double best = fabs(M[c * Vn + c]);
        for (int r = c + 1; r < Vn; r++) {
            const double v = fabs(M[r * Vn + c]);
            if (v > best) { best = v; piv = r; }
        }
        if (best == 0.0) {
            atomicOr(a.err_flag, ERR_SINGULAR);
            return;
        }
        if (piv != c) {
            for (int j = 0; j < Vn; j++) {
                const double tsw = M[c * Vn + j];
                M[c * Vn + j] = M[piv * Vn + j];
                M[piv * Vn + j] = tsw;
            }
            const double tsw = y[c];
            y[c] = y[piv];
            y[piv] = tsw;
        }
        const double inv = 1.0 / M[c * Vn + c];
        for (int r = c + 1; r < Vn; r++) {
            const double l = M[r * Vn + c] * inv;
            for (int j = c + 1; j < Vn; j++) M[r * Vn + j] -= l * M[c * Vn + j];
            y[r] -= l * y[c];
        }
    }
    for (int r = Vn - 1; r >= 0; r--) {
        double acc = y[r];
        for (int j = r + 1; j < Vn; j++) acc -= M[r * Vn + j] * y[j];
        y[r] = acc / M[r * Vn + r];
    }
    for (int i = 0; i < Vn; i++) a.pop[p * Vn + i] = y[i];
}

template <int VV>
__global__ void k_stateq(StateqArgs a)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.N) return;
    if (!a.m_active[a.cell_model[p]]) return;
    double M[VV * VV], y[VV];
    stateq_cell<VV>(a, p, VV, M, y);
}

constexpr int VMAX_GENERIC = 32;
constexpr int VMAX_BAND = 32;
__global__ void k_stateq_generic(StateqArgs a)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.N) return;
    if (!a.m_active[a.cell_model[p]]) return;
    double M[VMAX_GENERIC * VMAX_GENERIC], y[VMAX_GENERIC];
    stateq_cell<0>(a, p, a.V, M, y);
}

// ---- rate equations with band elements ----------------------------------------
// An element (p, k, nr) of Lambda* adds +v at (index(p, j_k), index(nr, i_k))
// and -v at (index(nr, i_k), index(nr, i_k)), v = -Ls opacity(p, k)
// (lineProducingSpecies.tpp:317-340): the second lies in cell nr's block, the
// first couples row p to cell nr.  The system is solved by block Jacobi: the
// block diagonal D (each cell's own triplets and the second kind of every
// element naming it) is factorised once, then each sweep solves
// D x' = y - O x for all cells from the last iterate x.  Sweeps stop when no
// population moves by more than BAND_TOL of itself or BAND_TOL_TOT of its
// cell's total (levels far below the total carry the rounding of the large
// ones and settle only to that); not converging in BAND_SWEEPS sweeps is an
// error.
constexpr int32_t ERR_BAND = 32;
constexpr double BAND_TOL = 1.0e-14;
constexpr double BAND_TOL_TOT = 1.0e-15;

struct BandArgs {
    int64_t S;
    const int32_t *cell;            // [N*S]
    const double  *Ls;              // [N*S*L]
    const int64_t *rev_off, *rev;   // slots naming each cell
    double  *lu;                    // [N*V*V]
    int32_t *piv;                   // [N*V]
    double  *x;                     // [N*V]
    int32_t *flags;                 // [BAND_SWEEPS + 1]
};

__device__ __forceinline__ double opacity_of(const StateqArgs &a, const double *pop, int k)
{
    return HH_OVER_FOUR_PI * (pop[a.jrad[k]] * a.Ba[k] - pop[a.irad[k]] * a.Bs[k]);
}

// History shift and factorisation of the cell's block of D (partial pivoting;
// multipliers below the diagonal, the pivot row of each column in piv).
template <int VV>
__device__ __forceinline__ void band_factor_cell(const StateqArgs &a, const BandArgs &b, int64_t p, int V,
                                                 double *M, double *y)
{
    const int Vn = (VV > 0) ? VV : V;
    stateq_assemble<VV>(a, p, Vn, M, y);
    // the diagonal terms of the elements of other cells that name this cell,
    // with their opacity from the populations before the solve
    for (int64_t e = b.rev_off[p]; e < b.rev_off[p + 1]; e++) {
        const int64_t sl = b.rev[e];
        const int64_t o = sl / b.S;
        const double *pop_o = a.pop + o * Vn;
        for (int k = 0; k < a.L; k++) {
            const int I = a.irad[k];
            if (I != Vn - 1) M[I * Vn + I] += b.Ls[sl * a.L + k] * opacity_of(a, pop_o, k);
        }
    }
    for (int c = 0; c < Vn; c++) {
        int pv = c;
        double best = fabs(M[c * Vn + c]);
        for (int r = c + 1; r < Vn; r++) {
            const double v = fabs(M[r * Vn + c]);
            if (v > best) { best = v; pv = r; }
        }
        if (best == 0.0) {
            atomicOr(a.err_flag, ERR_SINGULAR);
            pv = c;
        }
        b.piv[p * Vn + c] = pv;
        if (pv != c)
            for (int j = 0; j < Vn; j++) {
                const double tsw = M[c * Vn + j];
                M[c * Vn + j] = M[pv * Vn + j];
                M[pv * Vn + j] = tsw;
            }
        const double inv = best == 0.0 ? 0.0 : 1.0 / M[c * Vn + c];
        for (int r = c + 1; r < Vn; r++) {
            const double l = M[r * Vn + c] * inv;
            M[r * Vn + c] = l;
            for (int j = c + 1; j < Vn; j++) M[r * Vn + j] -= l * M[c * Vn + j];
        }
    }
    for (int i = 0; i < Vn * Vn; i++) b.lu[p * Vn * Vn + i] = M[i];
}

// One block-Jacobi sweep for the cell: right-hand side y - O x with x the
// current populations, solved with the stored factors into b.x.
template <int VV>
__device__ __forceinline__ void band_sweep_cell(const StateqArgs &a, const BandArgs &b, int64_t p, int V,
                                                double *M, double *y)
{
    const int Vn = (VV > 0) ? VV : V;
    for (int i = 0; i < Vn; i++) y[i] = 0.0;
    y[Vn - 1] = a.pop_tot[p];
    const double *pre = a.prev1 + p * Vn;   // the populations before the solve
    for (int64_t sl = p * b.S; sl < (p + 1) * b.S; sl++) {
        const int32_t c = b.cell[sl];
        if (c < 0) continue;
        for (int k = 0; k < a.L; k++) {
            const int Jl = a.jrad[k];
            if (Jl == Vn - 1) continue;
            const double v = -b.Ls[sl * a.L + k] * opacity_of(a, pre, k);
            y[Jl] -= v * a.pop[(int64_t)c * Vn + a.irad[k]];
        }
    }
    for (int i = 0; i < Vn * Vn; i++) M[i] = b.lu[p * Vn * Vn + i];
    // the factorisation swapped whole rows (multipliers included): every
    // interchange first, then the unit lower triangle
    for (int c = 0; c < Vn; c++) {
        const int pv = b.piv[p * Vn + c];
        if (pv != c) {
            const double tsw = y[c];
            y[c] = y[pv];
            y[pv] = tsw;
        }
    }
    for (int c = 0; c < Vn; c++)
        for (int r = c + 1; r < Vn; r++) y[r] -= M[r * Vn + c] * y[c];
    for (int r = Vn - 1; r >= 0; r--) {
        double acc = y[r];
        for (int j = r + 1; j < Vn; j++) acc -= M[r * Vn + j] * y[j];
        y[r] = acc / M[r * Vn + r];
    }
    for (int i = 0; i < Vn; i++) b.x[p * Vn + i] = y[i];
}

template <int VV>
__global__ void k_band_factor(StateqArgs a, BandArgs b)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.N) return;
    constexpr int SZ = VV > 0 ? VV : VMAX_BAND;
    double M[SZ * SZ], y[SZ];
    band_factor_cell<VV>(a, b, p, a.V, M, y);
}

template <int VV>
__global__ void k_band_sweep(StateqArgs a, BandArgs b, int t)
{
    if (b.flags[t] == 0) return;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.N) return;
    constexpr int SZ = VV > 0 ? VV : VMAX_BAND;
    double M[SZ * SZ], y[SZ];
    band_sweep_cell<VV>(a, b, p, a.V, M, y);
}

// Sweep t's iterate becomes the populations; a move above the tolerance
// asks for sweep t + 1.  change[t] keeps the sweep's largest move over the
// cell's total (as the bits of a non-negative double).
__global__ void k_band_commit(int64_t NV, int V, double *__restrict__ pop, const double *__restrict__ x,
                              const double *__restrict__ pop_tot, int32_t *__restrict__ flags,
                              unsigned long long *__restrict__ change, int t)
{
    if (flags[t] == 0) return;
    __shared__ double sh[256];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double d = 0.0;
    if (i < NV) {
        const double xn = x[i];
        const double mv = fabs(xn - pop[i]);
        const double tot = pop_tot[i / V];
        d = mv / tot;
        if (!(mv <= BAND_TOL * fabs(xn) || mv <= BAND_TOL_TOT * tot)) atomicOr(&flags[t + 1], 1);
        pop[i] = xn;
    }
    sh[threadIdx.x] = d;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0 && !(sh[0] <= 0.0)) atomicMax(&change[t], (unsigned long long)__double_as_longlong(sh[0]));
}

__global__ void k_band_flags(int32_t *flags, unsigned long long *change, int n)
{
    for (int i = threadIdx.x; i <= n; i += blockDim.x) {
        flags[i] = (i == 0);
        if (i < n) change[i] = 0ull;
    }
}

__global__ void k_band_check(const int32_t *flags, int n, int32_t *err_flag)
{
    if (flags[n]) atomicOr(err_flag, ERR_BAND);
}

__device__ __forceinline__ double block_sum(double v, double *sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

__device__ __forceinline__ double block_max(double v, double *sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// check_for_convergence (lineProducingSpecies.tpp:120-171): per block partial
// sums of the not-converged weight and the mean relative change, and the
// maximum relative change; levels with n <= 1e-10 n_tot are ignored.
// check_for_convergence (lineProducingSpecies.tpp:120-171) per model: each
// block reduces one chunk of cells of one model; the final kernel sums the
// chunks of every model in order (deterministic) and reports, for the batch,
// the worst model.
__global__ void k_conv_partial(int V, double pop_prec, const int64_t *__restrict__ chunk_cell,
                               const int32_t *__restrict__ cell_model, const int64_t *__restrict__ m_cell_off,
                               const double *__restrict__ pop, const double *__restrict__ prev1,
                               const double *__restrict__ pop_tot, double *__restrict__ partials,
                               int32_t *__restrict__ err_flag)
{
    __shared__ double sh[RED_BLOCK];
    const int64_t c0 = chunk_cell[blockIdx.x], c1 = chunk_cell[blockIdx.x + 1];
    const int32_t mdl = cell_model ? cell_model[c0] : 0;   // null: one model
    const double weight = 1.0 / (double)((m_cell_off[mdl + 1] - m_cell_off[mdl]) * V);
    double fnc = 0.0, rcm = 0.0, rmax = 0.0;
    bool finite = true;
    for (int64_t p = c0 + threadIdx.x; p < c1; p += blockDim.x) {
        const double min_pop = 1.0E-10 * pop_tot[p];
        for (int i = 0; i < V; i++) {
            const int64_t ind = p * V + i;
            // the criterion passes over NaN populations (the comparison below
            // is false for them), so a diverged iteration would look converged:
            // it is reported instead (err_flag null: not checked)
            finite = finite && __builtin_isfinite(pop[ind]);
            if (pop[ind] > min_pop) {
                double rc = 2.0;
                rc *= fabs(pop[ind] - prev1[ind]);
                rc /= (pop[ind] + prev1[ind]);
                if (rc > pop_prec) fnc += weight;
                rcm += weight * rc;
                rmax = fmax(rmax, rc);
            }
        }
    }
    if (!finite && err_flag) atomicOr(err_flag, ERR_NONFINITE);
    fnc = block_sum(fnc, sh);
    rcm = block_sum(rcm, sh);
    rmax = block_max(rmax, sh);
    if (threadIdx.x == 0) {
        partials[3 * blockIdx.x + 0] = fnc;
        partials[3 * blockIdx.x + 1] = rcm;
        partials[3 * blockIdx.x + 2] = rmax;
    }
}

__global__ void k_conv_final(int64_t M, const int64_t *__restrict__ m_chunk_off, const double *__restrict__ partials,
                             const int32_t *__restrict__ m_active, double *__restrict__ mconv,
                             double *__restrict__ out, const int32_t *__restrict__ err_flag)
{
    __shared__ double sh[RED_BLOCK];
    double a = 0.0, b = 0.0, c = 0.0;
    for (int64_t m = threadIdx.x; m < M; m += blockDim.x) {
        double fa = 0.0, fb = 0.0, fc = 0.0;
        for (int64_t i = m_chunk_off[m]; i < m_chunk_off[m + 1]; i++) {
            fa += partials[3 * i + 0];
            fb += partials[3 * i + 1];
            fc = fmax(fc, partials[3 * i + 2]);
        }
        mconv[4 * m + 0] = fa;
        mconv[4 * m + 1] = fb;
        mconv[4 * m + 2] = fc;
        mconv[4 * m + 3] = (double)m_active[m];
        if (m_active[m]) {
            a = fmax(a, fa);
            b = fmax(b, fb);
            c = fmax(c, fc);
        }
    }
    a = block_max(a, sh);
    b = block_max(b, sh);
    c = block_max(c, sh);
    if (threadIdx.x == 0) {
        out[0] = a;
        out[1] = b;
        out[2] = c;
        out[3] = (double)(*err_flag);
    }
}

// Ng acceleration (lineProducingSpecies.tpp:181-238) per model: five dot
// products over the model's populations, then x <- (1-a-b) x + a x1 + b x2
// and the history shift where the determinant is non-zero.
__global__ void k_ng_partial(int V, const int64_t *__restrict__ chunk_cell, const double *__restrict__ x,
                             const double *__restrict__ x1, const double *__restrict__ x2,
                             const double *__restrict__ x3, double *__restrict__ partials)
{
    __shared__ double sh[RED_BLOCK];
    const int64_t i0 = chunk_cell[blockIdx.x] * V, i1 = chunk_cell[blockIdx.x + 1] * V;
    double A1 = 0, A2 = 0, B2 = 0, C1 = 0, C2 = 0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        const double Q1 = x[i] - 2.0 * x1[i] + x2[i];
        const double Q2 = x[i] - x1[i] - x2[i] + x3[i];
        const double Q3 = x[i] - x1[i];
        A1 += Q1 * Q1; A2 += Q1 * Q2; B2 += Q2 * Q2; C1 += Q1 * Q3; C2 += Q2 * Q3;
    }
    A1 = block_sum(A1, sh); A2 = block_sum(A2, sh); B2 = block_sum(B2, sh);
    C1 = block_sum(C1, sh); C2 = block_sum(C2, sh);
    if (threadIdx.x == 0) {
        double *p = partials + 5 * blockIdx.x;
        p[0] = A1; p[1] = A2; p[2] = B2; p[3] = C1; p[4] = C2;
    }
}

__global__ void k_ng_final(int64_t M, const int64_t *__restrict__ m_chunk_off, const double *__restrict__ partials,
                           double *__restrict__ coef)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    double s[5] = {0, 0, 0, 0, 0};
    for (int64_t i = m_chunk_off[m]; i < m_chunk_off[m + 1]; i++)
        for (int q = 0; q < 5; q++) s[q] += partials[5 * i + q];
    const double A1 = s[0], A2 = s[1], B2 = s[2], C1 = s[3], C2 = s[4];
    const double B1 = A2;
    const double denominator = A1 * B2 - A2 * B1;
    coef[4 * m + 0] = denominator;
    coef[4 * m + 1] = (denominator != 0.0) ? (C1 * B2 - C2 * B1) / denominator : 0.0;
    coef[4 * m + 2] = (denominator != 0.0) ? (C2 * A1 - C1 * A2) / denominator : 0.0;
}

__global__ void k_ng_apply(int64_t N, int V, const int32_t *__restrict__ cell_model,
                           const int32_t *__restrict__ m_active, const double *__restrict__ coef,
                           double *__restrict__ x, double *__restrict__ x1, double *__restrict__ x2,
                           double *__restrict__ x3)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * V) return;
    const int32_t m = cell_model ? cell_model[i / V] : 0;   // null: one model
    if (!m_active[m] || coef[4 * m] == 0.0) return;
    const double a = coef[4 * m + 1], b = coef[4 * m + 2];
    const double xo = x[i];
    x[i] = (1.0 - a - b) * xo + a * x1[i] + b * x2[i];
    x3[i] = x2[i];
    x2[i] = x1[i];
    x1[i] = xo;
}

__global__ void k_zero_flag(int32_t *flag) { *flag = 0; }

// The driver's convergence words from the S species' words at 4 (l+1): the
// largest fraction, mean and maximum change; the error word is shared.  The
// context then holds one model, whose per-model words get the same figures.
__global__ void k_conv_species(int S, double *reduced, double *mconv)
{
    double a = 0.0, b = 0.0, c = 0.0;
    for (int l = 0; l < S; l++) {
        a = fmax(a, reduced[4 * (l + 1) + 0]);
        b = fmax(b, reduced[4 * (l + 1) + 1]);
        c = fmax(c, reduced[4 * (l + 1) + 2]);
    }
    reduced[0] = a;
    reduced[1] = b;
    reduced[2] = c;
    reduced[3] = reduced[4 * S + 3];
    mconv[0] = a;
    mconv[1] = b;
    mconv[2] = c;
}

// ---- Ng acceleration without allocation ---------------------------------------
// mgr_ng_acceleration_noalloc runs the kernels above on caller-owned arrays of
// one model.  Its scratch is one file-scope device array per GPU: the chunk
// partial sums (5 per chunk), the chunk bounds and the few per-model words
// the kernels read and write.  Every call overwrites what it reads first.
constexpr int64_t NOALLOC_MAX_CHUNKS = 4096;   // ncells <= NOALLOC_MAX_CHUNKS * CHUNK_CELLS
constexpr int64_t NA_PARTIALS = 0;                                     // double [5 * MAX_CHUNKS]
constexpr int64_t NA_CHUNK_CELL = NA_PARTIALS + 5 * NOALLOC_MAX_CHUNKS;   // int64 [MAX_CHUNKS + 1]
constexpr int64_t NA_WORDS = NA_CHUNK_CELL + NOALLOC_MAX_CHUNKS + 1;      // 8-byte words below
constexpr int64_t NA_M_CELL_OFF = NA_WORDS + 0, NA_M_CHUNK_OFF = NA_WORDS + 2, NA_COEF = NA_WORDS + 4,
                  NA_MCONV = NA_WORDS + 8, NA_REDUCED = NA_WORDS + 12, NA_FLAGS = NA_WORDS + 16;
constexpr int64_t NA_SIZE = NA_WORDS + 24;   // 8-byte words: 196,640 bytes
static __device__ __align__(256) double scratch[NA_SIZE];

// Chunk bounds of one model of ncells cells, its cell and chunk offsets, its
// activity flag (1) and a zero error word.
__global__ void k_noalloc_setup(int64_t ncells, int64_t nch, double *base)
{
    int64_t *chunk_cell = reinterpret_cast<int64_t *>(base + NA_CHUNK_CELL);
    for (int64_t c = threadIdx.x; c <= nch; c += blockDim.x) chunk_cell[c] = min(c * (int64_t)CHUNK_CELLS, ncells);
    if (threadIdx.x == 0) {
        int64_t *w = reinterpret_cast<int64_t *>(base);
        w[NA_M_CELL_OFF] = 0;
        w[NA_M_CELL_OFF + 1] = ncells;
        w[NA_M_CHUNK_OFF] = 0;
        w[NA_M_CHUNK_OFF + 1] = nch;
        int32_t *flags = reinterpret_cast<int32_t *>(base + NA_FLAGS);
        flags[0] = 1;   // m_active
        flags[1] = 0;   // err_flag
    }
}

unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }


}  // namespace

namespace mgr {

int launch_lambda(mgr_ctx *ctx)
{
    hipLaunchKernelGGL(k_lambda_sum, dim3(grid_for(ctx->N * ctx->L, 256)), dim3(256), 0, ctx->stream,
                       ctx->N, (int)ctx->L, (int)((ctx->F + WAVE - 1) / WAVE), ctx->Lpart, ctx->Ls);
    MGR_LAUNCH_CHECK();
    return MGR_OK;
}

int launch_jeff(mgr_ctx *ctx)
{
    for (int64_t l = 0; l < ctx->S; l++) {   // Simulation::compute_Jeff loops the species
        const int64_t Ll = ctx->sp_L[l], lo = ctx->sp_loff[l];
        hipLaunchKernelGGL(k_jeff, dim3(grid_for(ctx->N * Ll, 256)), dim3(256), 0, ctx->stream,
                           ctx->N, (int)Ll, (int)ctx->L, (int)lo, (int)ctx->Q, (int)ctx->F, (int)ctx->sp_V[l],
                           ctx->J, ctx->nr_line, ctx->qweights + l * ctx->Q, ctx->irad + lo,
                           ctx->pop + ctx->N * ctx->sp_voff[l], ctx->Ls, ctx->Jlin, ctx->Jeff, ctx->Jdif,
                           ctx->band_S, ctx->nband > 0 ? ctx->band_cell : nullptr, ctx->band_Ls);
        MGR_LAUNCH_CHECK();
    }
    return MGR_OK;
}

// The rate-equation arguments of species l.
static StateqArgs stateq_args(mgr_ctx *ctx, int64_t l)
{
    StateqArgs a;
    const int64_t lo = ctx->sp_loff[l], vo = ctx->N * ctx->sp_voff[l], c0 = ctx->sp_cp0[l];
    a.N = ctx->N; a.V = (int)ctx->sp_V[l]; a.L = (int)ctx->sp_L[l]; a.Lst = (int)ctx->L;
    a.nspecs = (int)ctx->nspecs; a.ncolpar = (int)ctx->sp_ncp[l];
    a.Jeff = ctx->Jeff + lo; a.Ls = ctx->Ls + lo; a.A = ctx->A + lo; a.Bs = ctx->Bs + lo; a.Ba = ctx->Ba + lo;
    a.abundance = ctx->abundance; a.T = ctx->temperature; a.pop_tot = ctx->pop_tot + ctx->N * l;
    a.irad = ctx->irad + lo; a.jrad = ctx->jrad + lo;
    a.cp_num = ctx->cp_num + c0; a.cp_ncol = ctx->cp_ncol + c0; a.cp_ntmp = ctx->cp_ntmp + c0;
    a.cp_koff = ctx->cp_koff + c0; a.cp_toff = ctx->cp_toff + c0; a.cp_coff = ctx->cp_coff + c0;
    a.cp_op = ctx->cp_op + c0;
    a.icol = ctx->icol; a.jcol = ctx->jcol; a.tmp = ctx->tmp; a.Cd = ctx->Cd; a.Ce = ctx->Ce;
    a.pop = ctx->pop + vo; a.prev1 = ctx->prev1 + vo; a.prev2 = ctx->prev2 + vo; a.prev3 = ctx->prev3 + vo;
    a.err_flag = ctx->err_flag;
    a.cell_model = ctx->d_cell_model;
    a.m_active = ctx->d_m_active;
    return a;
}

// The rate equations with band elements (one model): factorise the block
// diagonal, then BAND_SWEEPS block-Jacobi sweeps, each a no-op once an
// earlier one has converged (graph-capturable: the flags are set by kernels).
static int launch_stateq_band(mgr_ctx *ctx, const StateqArgs &a)
{
    if (ctx->V > VMAX_BAND) return ctx->fail(MGR_ERR_UNSUPPORTED, "more than 32 levels");
    BandArgs bb;
    bb.S = ctx->band_S;
    bb.cell = ctx->band_cell;
    bb.Ls = ctx->band_Ls;
    bb.rev_off = ctx->band_rev_off;
    bb.rev = ctx->band_rev;
    bb.lu = ctx->band_lu;
    bb.piv = ctx->band_piv;
    bb.x = ctx->band_x;
    bb.flags = ctx->band_flags;
    const dim3 g(grid_for(ctx->N, 128)), b(128);
    const int64_t NV = ctx->N * ctx->V;
    hipLaunchKernelGGL(k_band_flags, dim3(1), dim3(64), 0, ctx->stream, bb.flags, ctx->band_change, BAND_SWEEPS);
    MGR_LAUNCH_CHECK();
#define MGR_BAND_V(VV)                                                                                       \
    hipLaunchKernelGGL(k_band_factor<VV>, g, b, 0, ctx->stream, a, bb);                                       \
    for (int t = 0; t < BAND_SWEEPS; t++) {                                                                   \
        hipLaunchKernelGGL(k_band_sweep<VV>, g, b, 0, ctx->stream, a, bb, t);                                 \
        hipLaunchKernelGGL(k_band_commit, dim3(grid_for(NV, 256)), dim3(256), 0, ctx->stream, NV, (int)ctx->V, \
                           ctx->pop, (const double *)bb.x, (const double *)ctx->pop_tot, bb.flags,            \
                           ctx->band_change, t);                                                              \
    }
    switch (ctx->V) {
        case 2: MGR_BAND_V(2) break;
        case 3: MGR_BAND_V(3) break;
        case 4: MGR_BAND_V(4) break;
        case 5: MGR_BAND_V(5) break;
        case 6: MGR_BAND_V(6) break;
        case 7: MGR_BAND_V(7) break;
        case 8: MGR_BAND_V(8) break;
        default: MGR_BAND_V(0) break;
    }
#undef MGR_BAND_V
    MGR_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_band_check, dim3(1), dim3(1), 0, ctx->stream, (const int32_t *)bb.flags, BAND_SWEEPS,
                       ctx->err_flag);
    MGR_LAUNCH_CHECK();
    return MGR_OK;
}

static int launch_stateq_species(mgr_ctx *ctx, const StateqArgs &a);

int launch_stateq(mgr_ctx *ctx)
{
    if (ctx->nband > 0) return launch_stateq_band(ctx, stateq_args(ctx, 0));
    for (int64_t l = 0; l < ctx->S; l++) {   // Lines::iteration_using_statistical_equilibrium
        const int rc = launch_stateq_species(ctx, stateq_args(ctx, l));
        if (rc) return rc;
    }
    return MGR_OK;
}

static int launch_stateq_species(mgr_ctx *ctx, const StateqArgs &a)
{
    const dim3 g(grid_for(ctx->N, 128)), b(128);
    switch (a.V) {
        case 2: hipLaunchKernelGGL(k_stateq<2>, g, b, 0, ctx->stream, a); break;
        case 3: hipLaunchKernelGGL(k_stateq<3>, g, b, 0, ctx->stream, a); break;
        case 4: hipLaunchKernelGGL(k_stateq<4>, g, b, 0, ctx->stream, a); break;
        case 5: hipLaunchKernelGGL(k_stateq<5>, g, b, 0, ctx->stream, a); break;
        case 6: hipLaunchKernelGGL(k_stateq<6>, g, b, 0, ctx->stream, a); break;
        case 7: hipLaunchKernelGGL(k_stateq<7>, g, b, 0, ctx->stream, a); break;
        case 8: hipLaunchKernelGGL(k_stateq<8>, g, b, 0, ctx->stream, a); break;
        default:
            if (a.V > VMAX_GENERIC)
                return ctx->fail(MGR_ERR_UNSUPPORTED, "more than 32 levels");
            hipLaunchKernelGGL(k_stateq_generic, g, b, 0, ctx->stream, a);
    }
    MGR_LAUNCH_CHECK();
    return MGR_OK;
}

// check_for_convergence of every species; with several species each species'
// words go to reduced[4 (l+1) ..] and the driver's words are the largest over
// the species (the ALI goes on while any species has not converged,
// sim_lines.tpp:130-150).
int launch_convergence(mgr_ctx *ctx)
{
    for (int64_t l = 0; l < ctx->S; l++) {
        const int64_t vo = ctx->N * ctx->sp_voff[l];
        hipLaunchKernelGGL(k_conv_partial, dim3((unsigned)ctx->nchunks), dim3(RED_BLOCK), 0, ctx->stream,
                           (int)ctx->sp_V[l], ctx->pop_prec, ctx->d_chunk_cell, ctx->d_cell_model, ctx->d_m_cell_off,
                           ctx->pop + vo, ctx->prev1 + vo, ctx->pop_tot + ctx->N * l, ctx->partials, ctx->err_flag);
        MGR_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_conv_final, dim3(1), dim3(RED_BLOCK), 0, ctx->stream, ctx->M, ctx->d_m_chunk_off,
                           ctx->partials, ctx->d_m_active, ctx->d_mconv,
                           ctx->S > 1 ? ctx->reduced + 4 * (l + 1) : ctx->reduced, ctx->err_flag);
        MGR_LAUNCH_CHECK();
    }
    if (ctx->S > 1) {
        hipLaunchKernelGGL(k_conv_species, dim3(1), dim3(1), 0, ctx->stream, (int)ctx->S, ctx->reduced, ctx->d_mconv);
        MGR_LAUNCH_CHECK();
    }
    return MGR_OK;
}

// Ng acceleration of every species in turn (Lines::iteration_using_Ng_acceleration).
int launch_ng(mgr_ctx *ctx)
{
    for (int64_t l = 0; l < ctx->S; l++) {
        const int64_t vo = ctx->N * ctx->sp_voff[l], V = ctx->sp_V[l];
        hipLaunchKernelGGL(k_ng_partial, dim3((unsigned)ctx->nchunks), dim3(RED_BLOCK), 0, ctx->stream, (int)V,
                           ctx->d_chunk_cell, ctx->pop + vo, ctx->prev1 + vo, ctx->prev2 + vo, ctx->prev3 + vo,
                           ctx->partials);
        MGR_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_ng_final, dim3(grid_for(ctx->M, 64)), dim3(64), 0, ctx->stream, ctx->M,
                           ctx->d_m_chunk_off, ctx->partials, ctx->d_mng);
        MGR_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_ng_apply, dim3(grid_for(ctx->N * V, 256)), dim3(256), 0, ctx->stream, ctx->N,
                           (int)V, ctx->d_cell_model, ctx->d_m_active, ctx->d_mng, ctx->pop + vo, ctx->prev1 + vo,
                           ctx->prev2 + vo, ctx->prev3 + vo);
        MGR_LAUNCH_CHECK();
    }
    return MGR_OK;
}

// Enqueue the copy of the convergence words to pinned host memory.
static int enqueue_conv_copy(mgr_ctx *ctx)
{
    MGR_HIP(hipMemcpyAsync(ctx->h_reduced, ctx->reduced, 4 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    MGR_HIP(hipMemcpyAsync(ctx->h_mconv, ctx->d_mconv, ctx->M * 4 * sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream));
    return MGR_OK;
}

}  // namespace mgr

// Copy the host's per-model activity flags to the device (outside graphs).
int mgr::set_active(mgr_ctx *ctx, bool all)
{
    if (all)
        for (int64_t m = 0; m < ctx->M; m++) ctx->h_m_active[m] = 1;
    MGR_HIP(hipMemcpyAsync(ctx->d_m_active, ctx->h_m_active, ctx->M * sizeof(int32_t), hipMemcpyHostToDevice,
                           ctx->stream));
    return MGR_OK;
}

namespace mgr {

int read_convergence(mgr_ctx *ctx, mgr_conv *conv)
{
    MGR_HIP(hipStreamSynchronize(ctx->stream));
    const int32_t errv = (int32_t)ctx->h_reduced[3];
    if (conv) {
        conv->fraction_not_converged = ctx->h_reduced[0];
        conv->relative_change_mean = ctx->h_reduced[1];
        conv->relative_change_max = ctx->h_reduced[2];
    }
    if (errv & ERR_SINGULAR) return ctx->fail(MGR_ERR_NUMERIC, "singular rate matrix in a cell");
    if (errv & ERR_BAND)
        return ctx->fail(MGR_ERR_NUMERIC, "rate equations with Lambda* bands: block-Jacobi sweeps did not converge");
    if (errv & ERR_NONFINITE)
        return ctx->fail(MGR_ERR_NUMERIC, "non-finite level populations: the iteration diverged");
    if (errv) return ctx->fail(MGR_ERR_TRACE, "a ray pair ended off the boundary or an interior origin has an empty ray");
    return MGR_OK;
}

// One ALI iteration, enqueued on the context stream (graph-capturable: no
// memset nodes, the error word is cleared by a kernel).
static int enqueue_iteration(mgr_ctx *ctx, bool ng, int precision)
{
    int rc;
    hipLaunchKernelGGL(k_zero_flag, dim3(1), dim3(1), 0, ctx->stream, ctx->err_flag);
    MGR_LAUNCH_CHECK();
    if (ng) {
        if ((rc = launch_ng(ctx))) return rc;
    } else {
        if ((rc = launch_radiation_field(ctx, precision))) return rc;
        if ((rc = launch_jeff(ctx))) return rc;
        if ((rc = launch_stateq(ctx))) return rc;
    }
    if ((rc = launch_convergence(ctx))) return rc;
    if ((rc = launch_emissivity_opacity(ctx))) return rc;
    return enqueue_conv_copy(ctx);
}

static int build_graph(mgr_ctx *ctx, bool ng, int precision, hipGraphExec_t *out)
{
    hipGraph_t g = nullptr;
    MGR_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_iteration(ctx, ng, precision);
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return ctx->fail(MGR_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return ctx->fail(MGR_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
    return MGR_OK;
}

}  // namespace mgr

extern "C" {

int mgr_compute_Jeff(mgr_ctx *ctx)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    if (!ctx->J_set) return ctx->fail(MGR_ERR_STATE, "compute_radiation_field first");
    MGR_HIP(hipSetDevice(ctx->device));
    int rc = launch_jeff(ctx);
    if (rc) return rc;
    MGR_HIP(hipStreamSynchronize(ctx->stream));
    return MGR_OK;
}

int mgr_compute_level_populations_from_stateq(mgr_ctx *ctx, mgr_conv *conv)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    if (!ctx->J_set) return ctx->fail(MGR_ERR_STATE, "compute_radiation_field and compute_Jeff first");
    MGR_HIP(hipSetDevice(ctx->device));
    int rc;
    if ((rc = set_active(ctx, true))) return rc;
    hipLaunchKernelGGL(k_zero_flag, dim3(1), dim3(1), 0, ctx->stream, ctx->err_flag);
    MGR_LAUNCH_CHECK();
    if ((rc = launch_stateq(ctx))) return rc;
    if ((rc = launch_convergence(ctx))) return rc;
    if ((rc = launch_emissivity_opacity(ctx))) return rc;
    if ((rc = enqueue_conv_copy(ctx))) return rc;
    return read_convergence(ctx, conv);
}

int mgr_iteration_using_Ng_acceleration(mgr_ctx *ctx, mgr_conv *conv)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    if (!ctx->pops_set) return ctx->fail(MGR_ERR_STATE, "level populations not set");
    MGR_HIP(hipSetDevice(ctx->device));
    int rc;
    if ((rc = set_active(ctx, true))) return rc;
    hipLaunchKernelGGL(k_zero_flag, dim3(1), dim3(1), 0, ctx->stream, ctx->err_flag);
    MGR_LAUNCH_CHECK();
    if ((rc = launch_ng(ctx))) return rc;
    if ((rc = launch_convergence(ctx))) return rc;
    if ((rc = launch_emissivity_opacity(ctx))) return rc;
    if ((rc = enqueue_conv_copy(ctx))) return rc;
    return read_convergence(ctx, conv);
}

// Device address of this GPU's scratch array, looked up once per device
// with that device current.
static double *noalloc_scratch(int device)
{
    static std::mutex mtx;
    static double *addr[64] = {};
    if (device < 0 || device >= 64) return nullptr;
    std::lock_guard<std::mutex> lock(mtx);
    if (!addr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        addr[device] = static_cast<double *>(p);
    }
    return addr[device];
}

static int noalloc_fail(int code, const std::string &msg)
{
    mgr::set_context_free_error("mgr_ng_acceleration_noalloc: " + msg);
    return code;
}

int mgr_ng_acceleration_noalloc(int64_t ncells, int64_t nlev, double pop_prec, double *population,
                                double *prev1, double *prev2, double *prev3, const double *population_tot,
                                void *stream, mgr_conv *conv)
{
    if (ncells < 1 || nlev < 1) return noalloc_fail(MGR_ERR_ARGUMENT, "ncells and nlev must be positive");
    if (!population || !prev1 || !prev2 || !prev3 || !population_tot)
        return noalloc_fail(MGR_ERR_ARGUMENT, "null array");
    const int64_t nch = (ncells + CHUNK_CELLS - 1) / CHUNK_CELLS;
    if (nch > NOALLOC_MAX_CHUNKS)
        return noalloc_fail(MGR_ERR_ARGUMENT, "ncells " + std::to_string(ncells) + " exceeds the maximum of " +
                                                  std::to_string(NOALLOC_MAX_CHUNKS * CHUNK_CELLS));
    if (nlev > 4096) return noalloc_fail(MGR_ERR_ARGUMENT, "nlev exceeds 4096");
    int device = -1;
    if (hipGetDevice(&device) != hipSuccess) return noalloc_fail(MGR_ERR_DEVICE, "no current HIP device");
    double *base = noalloc_scratch(device);
    if (!base) return noalloc_fail(MGR_ERR_DEVICE, "scratch address of the current device not found");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const int V = (int)nlev;
    int64_t *chunk_cell = reinterpret_cast<int64_t *>(base + NA_CHUNK_CELL);
    int64_t *m_cell_off = reinterpret_cast<int64_t *>(base + NA_M_CELL_OFF);
    int64_t *m_chunk_off = reinterpret_cast<int64_t *>(base + NA_M_CHUNK_OFF);
    int32_t *m_active = reinterpret_cast<int32_t *>(base + NA_FLAGS);
    int32_t *err_flag = m_active + 1;
    double *partials = base + NA_PARTIALS, *coef = base + NA_COEF, *mconv = base + NA_MCONV,
           *reduced = base + NA_REDUCED;
    hipLaunchKernelGGL(k_noalloc_setup, dim3(1), dim3(256), 0, st, ncells, nch, base);
    // Lines::iteration_using_Ng_acceleration (lines.cpp:131-149) on one species:
    // update_using_Ng_acceleration, then check_for_convergence
    hipLaunchKernelGGL(k_ng_partial, dim3((unsigned)nch), dim3(RED_BLOCK), 0, st, V, chunk_cell, population, prev1,
                       prev2, prev3, partials);
    hipLaunchKernelGGL(k_ng_final, dim3(1), dim3(64), 0, st, (int64_t)1, m_chunk_off, partials, coef);
    hipLaunchKernelGGL(k_ng_apply, dim3(grid_for(ncells * nlev, 256)), dim3(256), 0, st, ncells, V, nullptr,
                       m_active, coef, population, prev1, prev2, prev3);
    hipLaunchKernelGGL(k_conv_partial, dim3((unsigned)nch), dim3(RED_BLOCK), 0, st, V, pop_prec, chunk_cell,
                       nullptr, m_cell_off, population, prev1, population_tot, partials, nullptr);
    hipLaunchKernelGGL(k_conv_final, dim3(1), dim3(RED_BLOCK), 0, st, (int64_t)1, m_chunk_off, partials, m_active,
                       mconv, reduced, err_flag);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return noalloc_fail(MGR_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
    if (conv) {
        double h[3];
        e = hipMemcpyAsync(h, reduced, sizeof(h), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return noalloc_fail(MGR_ERR_DEVICE, hipGetErrorString(e));
        conv->fraction_not_converged = h[0];
        conv->relative_change_mean = h[1];
        conv->relative_change_max = h[2];
    }
    return MGR_OK;
}

static int ensure_graphs(mgr_ctx *ctx, int precision)
{
    int rc;
    if (!ctx->traced && (rc = launch_trace(ctx))) return rc;
    if ((rc = set_active(ctx, true))) return rc;
    // recomputed in place when the frequencies changed; the graphs read them at replay
    if ((rc = launch_step_masks(ctx, precision == MGR_FP64))) return rc;
    if (ctx->nband > 0 && (rc = launch_step_masks(ctx, true))) return rc;   // the band kernel's
    if ((rc = launch_far_cache(ctx, precision))) return rc;
    if ((rc = tune_radiation_field(ctx, precision))) return rc;
    if (ctx->graph_precision != precision) {
        if (ctx->graph_normal) (void)hipGraphExecDestroy(ctx->graph_normal);
        if (ctx->graph_ng) (void)hipGraphExecDestroy(ctx->graph_ng);
        ctx->graph_normal = ctx->graph_ng = nullptr;
        ctx->graph_precision = -1;
        if ((rc = build_graph(ctx, false, precision, &ctx->graph_normal))) return rc;
        if ((rc = build_graph(ctx, true, precision, &ctx->graph_ng))) return rc;
        ctx->graph_precision = precision;
    }
    return MGR_OK;
}

int mgr_prepare(mgr_ctx *ctx, int precision)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    if (!ctx->spectral_set || ctx->spectral_image)
        return ctx->fail(MGR_ERR_STATE, "Spectral discretisation was not set for Lines");
    if (!ctx->bdy_set) return ctx->fail(MGR_ERR_STATE, "compute_boundary_intensities first");
    if (!ctx->pops_set) return ctx->fail(MGR_ERR_STATE, "level populations not set");
    if (precision != MGR_FP32 && precision != MGR_FP64) return ctx->fail(MGR_ERR_ARGUMENT, "bad precision");
    MGR_HIP(hipSetDevice(ctx->device));
    int rc = ensure_graphs(ctx, precision);
    if (rc) return rc;
    MGR_HIP(hipStreamSynchronize(ctx->stream));
    return MGR_OK;
}

int mgr_ali_iteration(mgr_ctx *ctx, int ng, int precision, mgr_conv *conv)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    if (!ctx->spectral_set || ctx->spectral_image)
        return ctx->fail(MGR_ERR_STATE, "Spectral discretisation was not set for Lines");
    if (!ctx->bdy_set) return ctx->fail(MGR_ERR_STATE, "compute_boundary_intensities first");
    if (!ctx->pops_set) return ctx->fail(MGR_ERR_STATE, "level populations not set");
    if (precision != MGR_FP32 && precision != MGR_FP64) return ctx->fail(MGR_ERR_ARGUMENT, "bad precision");
    MGR_HIP(hipSetDevice(ctx->device));
    int rc = ensure_graphs(ctx, precision);
    if (rc) return rc;
    if ((rc = set_active(ctx, true))) return rc;
    MGR_HIP(hipGraphLaunch(ng ? ctx->graph_ng : ctx->graph_normal, ctx->stream));
    rc = read_convergence(ctx, conv);
    if (rc == MGR_OK && !ng) ctx->J_set = true;
    return rc;
}

int mgr_compute_level_populations(mgr_ctx *ctx, int use_Ng_acceleration, int64_t max_niterations,
                                  int criterion, int precision, int64_t *iterations_out)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    if (iterations_out) *iterations_out = -1;
    // sim_lines.tpp:67-71: requires the line spectral discretisation
    if (!ctx->spectral_set || ctx->spectral_image)
        return ctx->fail(MGR_ERR_STATE, "Spectral discretisation was not set for Lines");
    if (!ctx->bdy_set) return ctx->fail(MGR_ERR_STATE, "compute_boundary_intensities first");
    if (!ctx->pops_set) return ctx->fail(MGR_ERR_STATE, "level populations not set");
    if (precision != MGR_FP32 && precision != MGR_FP64) return ctx->fail(MGR_ERR_ARGUMENT, "bad precision");
    if (criterion != MGR_CRIT_FRACTION && criterion != MGR_CRIT_MAX) return ctx->fail(MGR_ERR_ARGUMENT, "bad criterion");
    MGR_HIP(hipSetDevice(ctx->device));
    int rc = ensure_graphs(ctx, precision);
    if (rc) return rc;
    // Every model of a batch iterates until it meets the criterion (its own
    // iteration count, as if solved alone); converged models are frozen.
    if ((rc = set_active(ctx, true))) return rc;
    for (int64_t m = 0; m < ctx->M; m++) ctx->m_iterations[m] = 0;
    int64_t iteration = 0, iteration_normal = 0;
    bool some_not_converged = true;
    while (some_not_converged && iteration < max_niterations) {
        iteration++;
        const bool ng = use_Ng_acceleration && (iteration_normal == 4);
        MGR_HIP(hipGraphLaunch(ng ? ctx->graph_ng : ctx->graph_normal, ctx->stream));
        iteration_normal = ng ? 0 : iteration_normal + 1;
        if ((rc = read_convergence(ctx, nullptr))) return rc;
        some_not_converged = false;
        for (int64_t m = 0; m < ctx->M; m++) {
            if (!ctx->h_m_active[m]) continue;
            const mgr_conv c{ctx->h_mconv[4 * m + 0], ctx->h_mconv[4 * m + 1], ctx->h_mconv[4 * m + 2]};
            ctx->m_last_conv[m] = c;
            ctx->m_iterations[m] = iteration;
            const bool not_conv = (criterion == MGR_CRIT_MAX) ? c.relative_change_max >= ctx->pop_prec
                                                              : c.fraction_not_converged > 0.005;
            if (not_conv) some_not_converged = true;
            else ctx->h_m_active[m] = 0;
        }
        if (some_not_converged && ctx->M > 1 && (rc = set_active(ctx, false))) return rc;
    }
    ctx->J_set = true;
    if (iterations_out) *iterations_out = iteration;
    return MGR_OK;
}

}  // extern "C"
