// mgr_context.hip -- context life cycle, model upload, once-per-model set-up
// kernels (cell tables, spectral discretisation, boundary intensities, LTE)
// and the emissivity/opacity update, plus the data-movement entry points.
//
// Compiled with -ffp-contract=off: these kernels restate FP64 expressions of
// the reference and must round exactly as its scalar code does.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>

#include "mgr_common.hpp"

using namespace mgr;

namespace {
thread_local std::string g_create_error;
}  // namespace

namespace mgr {
// The message mgr_last_error(NULL) returns: failures of the entry points
// that take no context.
void set_context_free_error(const std::string &msg) { g_create_error = msg; }
}  // namespace mgr

namespace {

template <typename T>
int dalloc(mgr_ctx *ctx, T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return ctx->fail(MGR_ERR_NOMEM, std::string("hipMalloc of ") + std::to_string(n * sizeof(T)) +
                                            " bytes failed: " + hipGetErrorString(e));
    }
    return MGR_OK;
}

template <typename T>
int upload(mgr_ctx *ctx, T **p, const T *src, size_t n)
{
    int rc = dalloc(ctx, p, n);
    if (rc) return rc;
    if (n) MGR_HIP(hipMemcpy(*p, src, n * sizeof(T), hipMemcpyHostToDevice));
    return MGR_OK;
}

template <typename T>
int upload_cast(mgr_ctx *ctx, T **p, const int64_t *src, size_t n)
{
    std::vector<T> tmpv(n ? n : 1);
    for (size_t i = 0; i < n; i++) tmpv[i] = (T)src[i];
    return upload(ctx, p, tmpv.data(), n);
}

template <typename T>
void dfree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

// ---- kernels ---------------------------------------------------------------

// Thermodynamics::profile_width(inverse_mass, p) (thermodynamics.tpp:62-69) of
// every species, Simulation::get_dshift_max (simulation.tpp:15-37, the smallest
// over species), and the per-(cell, sorted line) window 11*width, 1/width used
// by get_eta_and_chi (simulation.tpp:403-446) with get_line_width's species.
__global__ void k_cell_tables(int64_t N, const LineTables *__restrict__ tb,
                              const double *__restrict__ T, const double *__restrict__ vturb2,
                              double max_width_fraction, double *__restrict__ pw,
                              double *__restrict__ pw_cmax, double *__restrict__ dshift_max,
                              double *__restrict__ W11, double *__restrict__ IW64, float *__restrict__ IW32)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const int L = tb->L;
    double dsm = DBL_MAX, wmax = 0.0;
    for (int l = 0; l < tb->S; l++) {
        const double w = sqrt(TWO_KB_OVER_AMU_C_SQUARED * tb->inv_mass[l] * T[p] + vturb2[p]);
        pw[l * N + p] = w;
        wmax = fmax(wmax, w);
        const double nd = max_width_fraction * w;
        if (dsm > nd) dsm = nd;
    }
    pw_cmax[p] = wmax;
    dshift_max[p] = dsm;
    for (int l = 0; l < L; l++) {
        const double width = tb->line[l] * pw[tb->width_spec[l] * N + p];
        W11[p * L + l] = LINE_WINDOW * width;
        const double iw = 1.0 / width;
        IW64[p * L + l] = iw;
        IW32[p * L + l] = (float)iw;
    }
}

// Simulation::compute_spectral_discretisation (sim_spectrum.tpp:13-196):
// nu(p, i) = line[g] + line[g]*pw(s, p)*root(s, z) for i = g*Q + z, where g
// runs over the lines species after species, line[g] is the sorted list at
// that running index and s is the species of line g; sorted per cell.  With
// pw null, compute_spectral_discretisation_image (sim_spectrum.tpp:206-311):
// the same grid with the one width `image_width` in every cell.
// One thread per (p, i) ranks its value among the cell's F values (ties by
// original index), which gives heapsort's order for distinct values.
__device__ __forceinline__ double bin_value(const LineTables *__restrict__ tb, const double *__restrict__ pw,
                                            int64_t N, int64_t p, double image_width,
                                            const double *__restrict__ roots, int i)
{
    const int Q = tb->Q;
    const int g = i / Q, z = i % Q;
    const int s = tb->line_spec[g];
    const double w = pw ? pw[s * N + p] : image_width;
    const double lg = tb->line[g];
    return lg + (lg * w) * roots[s * Q + z];
}

// The same with one workgroup per cell: the cell's F values are formed once
// into LDS and each ranked against them there (F doubles of LDS).
__global__ void k_spectral_cell(int64_t N, const LineTables *__restrict__ tb,
                                const double *__restrict__ pw, double image_width,
                                const double *__restrict__ roots, double *__restrict__ nu,
                                int32_t *__restrict__ nr_line)
{
    extern __shared__ double vals[];
    const int F = tb->F, Q = tb->Q, L = tb->L;
    const int64_t p = blockIdx.x;
    for (int i = threadIdx.x; i < F; i += blockDim.x) vals[i] = bin_value(tb, pw, N, p, image_width, roots, i);
    __syncthreads();
    for (int i = threadIdx.x; i < F; i += blockDim.x) {
        const double vi = vals[i];
        int rank = 0;
        for (int j = 0; j < F; j++) {
            const double vj = vals[j];
            rank += (vj < vi) || (vj == vi && j < i);
        }
        nu[p * F + rank] = vi;
        nr_line[(p * L + i / Q) * Q + i % Q] = rank;
    }
}

__global__ void k_spectral(int64_t N, const LineTables *__restrict__ tb,
                           const double *__restrict__ pw, double image_width,
                           const double *__restrict__ roots, double *__restrict__ nu,
                           int32_t *__restrict__ nr_line)
{
    const int F = tb->F, Q = tb->Q, L = tb->L;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * F) return;
    const int64_t p = t / F;
    const int i = (int)(t % F);
    const int ki = i / Q, zi = i % Q;
    const double vi = bin_value(tb, pw, N, p, image_width, roots, i);
    int rank = 0;
    for (int j = 0; j < F; j++) {
        const double vj = bin_value(tb, pw, N, p, image_width, roots, j);
        rank += (vj < vi) || (vj == vi && j < i);
    }
    nu[p * F + rank] = vi;
    nr_line[(p * L + ki) * Q + zi] = rank;
    (void)L;
}

// Simulation::compute_boundary_intensities (sim_radiation.tpp:17-48) with
// planck (src/Functions/planck.tpp:15-24).
__global__ void k_boundary_intensities(int64_t NB, int F, const int32_t *__restrict__ bdy2cell,
                                       const double *__restrict__ temps, double T_default,
                                       const double *__restrict__ nu, double *__restrict__ I_bdy)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NB * F) return;
    const int64_t b = t / F;
    const int f = (int)(t % F);
    const int64_t p = bdy2cell[b];
    const double T = temps ? temps[b] : T_default;
    const double freq = nu[p * F + f];
    const double h_over_kbT = HH_OVER_KB / T;
    I_bdy[t] = TWO_HH_OVER_CC_SQUARED * (freq * freq * freq) / expm1(h_over_kbT * freq);
}

// LineProducingSpecies::update_using_LTE (lineProducingSpecies.tpp:82-115).
__global__ void k_lte(int64_t N, int V, int nspecs, int num, const double *__restrict__ abundance,
                      const double *__restrict__ T, const double *__restrict__ weight,
                      const double *__restrict__ energy, double *__restrict__ pop,
                      double *__restrict__ pop_tot)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const double ptot = abundance[p * nspecs + num];
    pop_tot[p] = ptot;
    double pf = 0.0;
    for (int i = 0; i < V; i++) {
        const double v = weight[i] * exp(-energy[i] / (KB * T[p]));
        pop[p * V + i] = v;
        pf += v;
    }
    for (int i = 0; i < V; i++) pop[p * V + i] *= ptot / pf;
}

__global__ void k_pop_tot(int64_t N, int nspecs, int num, const double *__restrict__ abundance,
                          double *__restrict__ pop_tot)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    pop_tot[p] = abundance[p * nspecs + num];
}

}  // namespace

// LineProducingSpecies::get_emissivity / get_opacity (lineProducingSpecies.tpp:41-70),
// Lines::set_emissivity_and_opacity (lines.tpp:44-60), and the scaled line
// prefactors of the radiation-field kernel: eta = sum emissivity nu phi with
// phi = (1/width)(1/sqrt(pi)) exp(-x^2), so E = emissivity/(width sqrt(pi))
// scaled by CHI_SCALE*I_SCALE, and K = opacity/(width sqrt(pi)) * CHI_SCALE,
// both stored at the sorted line position l.
__global__ void k_emissivity_opacity(int64_t N, const LineTables *__restrict__ tb,
                                     const double *__restrict__ A, const double *__restrict__ Bs,
                                     const double *__restrict__ Ba, const int32_t *__restrict__ irad,
                                     const int32_t *__restrict__ jrad, const double *__restrict__ pop_all,
                                     const double *__restrict__ IW64, const double *__restrict__ W11,
                                     double *__restrict__ emis, double *__restrict__ opac,
                                     float4 *__restrict__ row32, double4 *__restrict__ row64)
{
    const int L = tb->L;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * L) return;
    const int64_t p = t / L;
    const int l = (int)(t % L);
    const int k = tb->line_index[l];
    // the line's species: its populations, levels i, j of its own numbering
    const int sp = tb->line_spec[k];
    const int V = tb->sV[sp];
    const double *pop = pop_all + (int64_t)tb->voff[sp] * N;
    const int i = irad[k], j = jrad[k];
    const double e = HH_OVER_FOUR_PI * A[k] * pop[p * V + i];
    const double o = HH_OVER_FOUR_PI * (pop[p * V + j] * Ba[k] - pop[p * V + i] * Bs[k]);
    emis[p * L + k] = e;
    opac[p * L + k] = o;
    const double pref = INVERSE_SQRT_PI * IW64[p * L + l];
    const double es = e * (CHI_SCALE * I_SCALE) * pref;
    const double ks = o * CHI_SCALE * pref;
    const double iw = IW64[p * L + l];
    row64[p * L + l] = make_double4(W11[p * L + l], iw, es, ks);
    // FP32 row: 1/width scaled by sqrt(log2 e) so that exp(-x^2) = exp2(-x'^2)
    row32[p * L + l] = make_float4((float)(iw * SQRT_LOG2E), (float)es, (float)ks, 0.0f);
}

namespace mgr {

static unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

int launch_cell_tables(mgr_ctx *ctx)
{
    hipLaunchKernelGGL(k_cell_tables, dim3(grid_for(ctx->N, 256)), dim3(256), 0, ctx->stream,
                       ctx->N, ctx->tables, ctx->temperature, ctx->vturb2, ctx->max_width_fraction,
                       ctx->pw, ctx->pw_cmax, ctx->dshift_max, ctx->W11, ctx->IW64, ctx->IW32);
    MGR_LAUNCH_CHECK();
    return MGR_OK;
}

int launch_emissivity_opacity(mgr_ctx *ctx)
{
    hipLaunchKernelGGL(k_emissivity_opacity, dim3(grid_for(ctx->N * ctx->L, 256)), dim3(256), 0,
                       ctx->stream, ctx->N, ctx->tables, ctx->A, ctx->Bs, ctx->Ba, ctx->irad,
                       ctx->jrad, ctx->pop, ctx->IW64, ctx->W11, ctx->emis, ctx->opac, ctx->row32,
                       ctx->row64);
    MGR_LAUNCH_CHECK();
    return MGR_OK;
}

// Validation of one model view on its own (sizes and index ranges).
static int check_view(mgr_ctx *ctx, const mgr_model_view *m, int64_t idx)
{
    const std::string who = "model " + std::to_string(idx) + ": ";
    if (m->ncells <= 0 || m->nrays <= 0 || (m->nrays % 2) != 0 || m->nrad <= 0 || m->nlev < 2 ||
        m->nquads <= 0 || m->nboundary < 0)
        return ctx->fail(MGR_ERR_ARGUMENT, who + "model sizes out of range (need ncells, nrad, nquads > 0, even nrays, nlev >= 2)");
    if (!m->position || !m->velocity || !m->n_neighbors || !m->rays || !m->weights || !m->temperature ||
        !m->vturb2 || !m->abundance || !m->energy || !m->weight || !m->irad || !m->jrad || !m->A || !m->Bs ||
        !m->Ba || !m->frequency || !m->roots || !m->qweights || (m->nboundary > 0 && !m->boundary2cell_nr))
        return ctx->fail(MGR_ERR_ARGUMENT, who + "null array in the model view");
    for (int64_t p = 0; p < m->ncells; p++)
        if (m->n_neighbors[p] < 0) return ctx->fail(MGR_ERR_ARGUMENT, who + "negative n_neighbors");
    int64_t M = 0;
    for (int64_t p = 0; p < m->ncells; p++) M += m->n_neighbors[p];
    if (M > 0 && !m->neighbors) return ctx->fail(MGR_ERR_ARGUMENT, who + "null neighbours");
    for (int64_t q = 0; q < M; q++)
        if (m->neighbors[q] < 0 || m->neighbors[q] >= m->ncells)
            return ctx->fail(MGR_ERR_ARGUMENT, who + "neighbour index out of range");
    for (int64_t b = 0; b < m->nboundary; b++)
        if (m->boundary2cell_nr[b] < 0 || m->boundary2cell_nr[b] >= m->ncells)
            return ctx->fail(MGR_ERR_ARGUMENT, who + "boundary2cell_nr out of range");
    if (m->nlspecs > 1) {
        if (!m->more_species) return ctx->fail(MGR_ERR_ARGUMENT, who + "nlspecs > 1 with no more_species");
        for (int64_t l = 1; l < m->nlspecs; l++) {
            const mgr_species_view *v = &m->more_species[l - 1];
            if (v->nrad <= 0 || v->nlev < 2)
                return ctx->fail(MGR_ERR_ARGUMENT, who + "species " + std::to_string(l) + ": sizes out of range");
            if (!v->energy || !v->weight || !v->irad || !v->jrad || !v->A || !v->Bs || !v->Ba || !v->frequency ||
                !v->roots || !v->qweights)
                return ctx->fail(MGR_ERR_ARGUMENT, who + "species " + std::to_string(l) + ": null array");
        }
    }
    return MGR_OK;
}

// The species of a model view: species 0 from its own fields, then more_species.
static std::vector<mgr_species_view> species_of(const mgr_model_view *m)
{
    std::vector<mgr_species_view> v(1);
    mgr_species_view &s = v[0];
    s.num = m->num; s.inverse_mass = m->inverse_mass; s.nlev = m->nlev; s.nrad = m->nrad;
    s.energy = m->energy; s.weight = m->weight; s.irad = m->irad; s.jrad = m->jrad;
    s.A = m->A; s.Bs = m->Bs; s.Ba = m->Ba; s.frequency = m->frequency;
    s.ncolpar = m->ncolpar; s.colpar_num = m->colpar_num; s.colpar_orth_or_para = m->colpar_orth_or_para;
    s.colpar_ncol = m->colpar_ncol; s.colpar_ntmp = m->colpar_ntmp; s.icol = m->icol; s.jcol = m->jcol;
    s.tmp = m->tmp; s.Cd = m->Cd; s.Ce = m->Ce; s.roots = m->roots; s.qweights = m->qweights;
    for (int64_t l = 1; l < m->nlspecs; l++) v.push_back(m->more_species[l - 1]);
    return v;
}

template <typename T>
static bool same(const T *a, const T *b, int64_t n)
{
    if (n <= 0) return true;
    if (!a || !b) return a == b;
    return std::memcmp(a, b, (size_t)n * sizeof(T)) == 0;
}

// Models of one batch must share the species, line data, collision tables,
// quadrature and iteration parameters (only the geometry, the thermodynamics
// and the abundances differ).
static int check_same_physics(mgr_ctx *ctx, const mgr_model_view *a, const mgr_model_view *b, int64_t idx)
{
    const std::string who = "model " + std::to_string(idx) + ": ";
    bool ok = a->nspecs == b->nspecs && a->nquads == b->nquads && a->num == b->num && a->nlev == b->nlev &&
              a->nrad == b->nrad && a->ncolpar == b->ncolpar && a->pop_prec == b->pop_prec &&
              a->max_width_fraction == b->max_width_fraction && a->n_off_diag == b->n_off_diag &&
              a->inverse_mass == b->inverse_mass;
    if (ok) {
        const int64_t V = a->nlev, L = a->nrad, Q = a->nquads;
        ok = same(a->energy, b->energy, V) && same(a->weight, b->weight, V) && same(a->irad, b->irad, L) &&
             same(a->jrad, b->jrad, L) && same(a->A, b->A, L) && same(a->Bs, b->Bs, L) && same(a->Ba, b->Ba, L) &&
             same(a->frequency, b->frequency, L) && same(a->roots, b->roots, Q) &&
             same(a->qweights, b->qweights, Q) && same(a->colpar_num, b->colpar_num, a->ncolpar) &&
             same(a->colpar_ncol, b->colpar_ncol, a->ncolpar) && same(a->colpar_ntmp, b->colpar_ntmp, a->ncolpar) &&
             same(a->colpar_orth_or_para, b->colpar_orth_or_para, a->ncolpar);
        if (ok) {
            int64_t nk = 0, nt = 0, nc = 0;
            for (int64_t c = 0; c < a->ncolpar; c++) {
                nk += a->colpar_ncol[c];
                nt += a->colpar_ntmp[c];
                nc += a->colpar_ncol[c] * a->colpar_ntmp[c];
            }
            ok = same(a->icol, b->icol, nk) && same(a->jcol, b->jcol, nk) && same(a->tmp, b->tmp, nt) &&
                 same(a->Cd, b->Cd, nc) && same(a->Ce, b->Ce, nc);
        }
    }
    if (!ok) return ctx->fail(MGR_ERR_ARGUMENT, who + "a batch needs the same species, line data and quadrature in every model");
    return MGR_OK;
}

int upload_models(mgr_ctx *ctx, const mgr_model_view *models, int64_t nmodels)
{
    int rc;
    if (nmodels < 1) return ctx->fail(MGR_ERR_ARGUMENT, "a batch needs at least one model");
    for (int64_t k = 0; k < nmodels; k++) {
        if ((rc = check_view(ctx, &models[k], k))) return rc;
        if (k > 0 && (rc = check_same_physics(ctx, &models[0], &models[k], k))) return rc;
    }
    const mgr_model_view *m0 = &models[0];
    const int64_t Mm = nmodels;
    ctx->M = Mm;

    // merged geometry, thermodynamics and abundances (cell indices offset per model)
    ctx->m_cell_off.assign(Mm + 1, 0);
    int64_t nb_tot = 0, nbd_tot = 0;
    for (int64_t k = 0; k < Mm; k++) {
        ctx->m_cell_off[k + 1] = ctx->m_cell_off[k] + models[k].ncells;
        for (int64_t p = 0; p < models[k].ncells; p++) nb_tot += models[k].n_neighbors[p];
        nbd_tot += models[k].nboundary;
    }
    const int64_t N = ctx->m_cell_off[Mm];
    if (N > INT32_MAX / 2) return ctx->fail(MGR_ERR_UNSUPPORTED, "ncells exceeds the 32-bit cell index");
    std::vector<double> pos, vel, temp, vt2, abund;
    std::vector<int64_t> nnb, nbs, b2c;
    pos.reserve(3 * N); vel.reserve(3 * N); temp.reserve(N); vt2.reserve(N);
    abund.reserve(N * m0->nspecs); nnb.reserve(N); nbs.reserve(nb_tot); b2c.reserve(nbd_tot);
    // distinct ray lists become ray sets
    std::vector<double> rays, rw;
    std::vector<int64_t> set_off{0};
    std::vector<int64_t> model_set(Mm);
    for (int64_t k = 0; k < Mm; k++) {
        const mgr_model_view &v = models[k];
        const int64_t c0 = ctx->m_cell_off[k];
        pos.insert(pos.end(), v.position, v.position + 3 * v.ncells);
        vel.insert(vel.end(), v.velocity, v.velocity + 3 * v.ncells);
        temp.insert(temp.end(), v.temperature, v.temperature + v.ncells);
        vt2.insert(vt2.end(), v.vturb2, v.vturb2 + v.ncells);
        abund.insert(abund.end(), v.abundance, v.abundance + v.ncells * v.nspecs);
        int64_t e = 0;
        for (int64_t p = 0; p < v.ncells; p++) {
            nnb.push_back(v.n_neighbors[p]);
            for (int64_t j = 0; j < v.n_neighbors[p]; j++) nbs.push_back(v.neighbors[e + j] + c0);
            e += v.n_neighbors[p];
        }
        for (int64_t b = 0; b < v.nboundary; b++) b2c.push_back(v.boundary2cell_nr[b] + c0);
        int64_t found = -1;
        for (size_t sidx = 0; sidx + 1 < set_off.size() && found < 0; sidx++) {
            const int64_t R0 = set_off[sidx + 1] - set_off[sidx];
            if (R0 == v.nrays && same(rays.data() + 3 * set_off[sidx], v.rays, 3 * R0) &&
                same(rw.data() + set_off[sidx], v.weights, R0))
                found = (int64_t)sidx;
        }
        if (found < 0) {
            found = (int64_t)set_off.size() - 1;
            rays.insert(rays.end(), v.rays, v.rays + 3 * v.nrays);
            rw.insert(rw.end(), v.weights, v.weights + v.nrays);
            set_off.push_back(set_off.back() + v.nrays);
        }
        model_set[k] = found;
    }
    const int64_t Rtot = set_off.back();

    mgr_model_view mv = *m0;
    mv.ncells = N;
    mv.nboundary = nbd_tot;
    mv.nrays = Rtot;
    mv.position = pos.data();
    mv.velocity = vel.data();
    mv.n_neighbors = nnb.data();
    mv.neighbors = nbs.data();
    mv.rays = rays.data();
    mv.weights = rw.data();
    mv.boundary2cell_nr = b2c.data();
    mv.temperature = temp.data();
    mv.vturb2 = vt2.data();
    mv.abundance = abund.data();
    const mgr_model_view *m = &mv;

    // per-model ray sets and (origin, ray) offsets
    ctx->m_R.assign(Mm, 0);
    ctx->m_rbase.assign(Mm, 0);
    ctx->m_q_off.assign(Mm + 1, 0);
    ctx->h_cell_qbase.assign(N, 0);
    std::vector<int32_t> cell_model(N);
    int64_t Rmax = 0;
    for (int64_t k = 0; k < Mm; k++) {
        const int64_t sidx = model_set[k];
        ctx->m_R[k] = (int32_t)(set_off[sidx + 1] - set_off[sidx]);
        ctx->m_rbase[k] = (int32_t)set_off[sidx];
        Rmax = std::max<int64_t>(Rmax, ctx->m_R[k]);
        ctx->m_q_off[k + 1] = ctx->m_q_off[k] + models[k].ncells * ctx->m_R[k];
        for (int64_t p = ctx->m_cell_off[k]; p < ctx->m_cell_off[k + 1]; p++) {
            cell_model[p] = (int32_t)k;
            ctx->h_cell_qbase[p] = ctx->m_q_off[k] + (p - ctx->m_cell_off[k]) * ctx->m_R[k];
        }
    }
    ctx->NQ = ctx->m_q_off[Mm];

    // reduction chunks: CHUNK_CELLS cells of one model each
    ctx->m_chunk_off.assign(Mm + 1, 0);
    std::vector<int64_t> chunk_cell{0};
    for (int64_t k = 0; k < Mm; k++) {
        for (int64_t c = ctx->m_cell_off[k]; c < ctx->m_cell_off[k + 1]; c += CHUNK_CELLS)
            chunk_cell.push_back(std::min<int64_t>(c + CHUNK_CELLS, ctx->m_cell_off[k + 1]));
        ctx->m_chunk_off[k + 1] = (int64_t)chunk_cell.size() - 1;
    }
    ctx->nchunks = (int64_t)chunk_cell.size() - 1;

    // the line-producing species (Lines::setup, lines.cpp:43-87): lines and
    // levels species after species
    const std::vector<mgr_species_view> spv = species_of(m);
    const int64_t S = (int64_t)spv.size();
    if (S > MAX_SPECIES) return ctx->fail(MGR_ERR_UNSUPPORTED, "more than " + std::to_string(MAX_SPECIES) + " line-producing species");
    if (S > 1 && Mm > 1) return ctx->fail(MGR_ERR_UNSUPPORTED, "several line-producing species in a batch of models are not supported");
    if (S > 1 && m->n_off_diag >= 2)
        return ctx->fail(MGR_ERR_UNSUPPORTED, "n_off_diag >= 2 (banded ALO) with several line-producing species is not supported");
    ctx->S = S;
    ctx->sp_V.assign(S, 0); ctx->sp_L.assign(S, 0); ctx->sp_voff.assign(S, 0); ctx->sp_loff.assign(S, 0);
    ctx->sp_num.assign(S, 0); ctx->sp_cp0.assign(S, 0); ctx->sp_ncp.assign(S, 0);
    int64_t Vt = 0, Lt = 0, Ct = 0;
    for (int64_t l = 0; l < S; l++) {
        ctx->sp_V[l] = spv[l].nlev; ctx->sp_L[l] = spv[l].nrad; ctx->sp_voff[l] = Vt; ctx->sp_loff[l] = Lt;
        ctx->sp_num[l] = spv[l].num; ctx->sp_cp0[l] = Ct; ctx->sp_ncp[l] = spv[l].ncolpar;
        Vt += spv[l].nlev; Lt += spv[l].nrad; Ct += spv[l].ncolpar;
        if (spv[l].num < 0 || spv[l].num >= m->nspecs) return ctx->fail(MGR_ERR_ARGUMENT, "line species index out of range");
    }

    ctx->N = N; ctx->R = Rmax; ctx->NB = nbd_tot; ctx->nspecs = m->nspecs;
    ctx->L = Lt; ctx->V = Vt; ctx->Q = m->nquads; ctx->F = Lt * m->nquads;
    ctx->pop_prec = m->pop_prec; ctx->max_width_fraction = m->max_width_fraction;
    ctx->num = m->num; ctx->ncolpar = Ct;
    const int64_t R = Rtot, NB = ctx->NB, L = ctx->L, V = ctx->V, Q = ctx->Q;

    if (L > MAX_LINES) return ctx->fail(MGR_ERR_UNSUPPORTED, "more than 64 line transitions");
    // n_off_diag = 1 is the diagonal operator: RayPair::update_Lambda adds
    // bands m = 1 .. n_off_diag-1 only (raypair_lambda.tpp:129-146), and the
    // back-substitution branch's L_diag (raypair_solver.tpp:358-370) is the
    // same diagonal of the inverse.  Wider bands couple the cells' rate
    // equations; they are built for one model at a time.
    if (m->n_off_diag < 0 || m->n_off_diag > MAX_OFF_DIAG)
        return ctx->fail(MGR_ERR_UNSUPPORTED, "n_off_diag above " + std::to_string(MAX_OFF_DIAG) + " is not supported");
    if (m->n_off_diag >= 2 && Mm > 1)
        return ctx->fail(MGR_ERR_UNSUPPORTED, "n_off_diag >= 2 (banded ALO) in a batch of models is not supported");
    ctx->nband = m->n_off_diag >= 2 ? (int)m->n_off_diag - 1 : 0;
    ctx->band_S = (ctx->R / 2) * 2 * ctx->nband;
    // Lines::setup (src/Model/Lines/lines.cpp:43-87): sorted line list.
    LineTables &tb = ctx->h_tables;
    std::memset(&tb, 0, sizeof(tb));
    tb.L = (int)L; tb.V = (int)V; tb.Q = (int)Q; tb.F = (int)ctx->F; tb.S = (int)S;
    tb.inverse_mass = m->inverse_mass;
    std::vector<double> freq_all(L), A_all(L), Bs_all(L), Ba_all(L), energy_all(V), weight_all(V);
    std::vector<double> roots_all(S * Q), qw_all(S * Q);
    std::vector<int64_t> irad_all(L), jrad_all(L);
    for (int64_t l = 0; l < S; l++) {
        const mgr_species_view &v = spv[l];
        tb.sV[l] = (int)v.nlev; tb.sL[l] = (int)v.nrad;
        tb.voff[l] = (int)ctx->sp_voff[l]; tb.loff[l] = (int)ctx->sp_loff[l];
        tb.inv_mass[l] = v.inverse_mass;
        for (int64_t k = 0; k < v.nrad; k++) {
            const int64_t g = ctx->sp_loff[l] + k;
            if (v.irad[k] < 0 || v.irad[k] >= v.nlev || v.jrad[k] < 0 || v.jrad[k] >= v.nlev)
                return ctx->fail(MGR_ERR_ARGUMENT, "radiative transition level out of range");
            tb.line_spec[g] = (int)l;
            freq_all[g] = v.frequency[k]; A_all[g] = v.A[k]; Bs_all[g] = v.Bs[k]; Ba_all[g] = v.Ba[k];
            irad_all[g] = v.irad[k]; jrad_all[g] = v.jrad[k];
        }
        for (int64_t i = 0; i < v.nlev; i++) {
            energy_all[ctx->sp_voff[l] + i] = v.energy[i];
            weight_all[ctx->sp_voff[l] + i] = v.weight[i];
        }
        for (int64_t z = 0; z < Q; z++) { roots_all[l * Q + z] = v.roots[z]; qw_all[l * Q + z] = v.qweights[z]; }
    }
    {
        std::vector<int> order(L);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return freq_all[a] < freq_all[b]; });
        for (int l = 0; l < L; l++) { tb.line[l] = freq_all[order[l]]; tb.line_index[l] = order[l]; }
        // get_line_width: corresponding_l_for_spec[line_index[l]], with the
        // back-map laid out over bins (species, transition, root) in order
        for (int l = 0; l < L; l++) {
            const int64_t f = tb.line_index[l];
            int sw = 0;
            for (int64_t q = 0; q < S; q++)
                if (f >= ctx->sp_loff[q] * Q) sw = (int)q;
            tb.width_spec[l] = sw;
        }
    }

    // neighbours (CSR); every index was checked against its own model above,
    // so a ray never leaves its model: the tracer trusts these indices
    std::vector<int64_t> off(N + 1, 0);
    for (int64_t p = 0; p < N; p++) off[p + 1] = off[p] + m->n_neighbors[p];
    const int64_t M = off[N];

    // Boundary::read (boundary.cpp:19-51)
    std::vector<int32_t> c2b(N, -1), isb_i(N, 0);
    std::vector<uint8_t> isb(N, 0);
    ctx->boundary2cell.assign(m->boundary2cell_nr, m->boundary2cell_nr + NB);
    for (int64_t b = 0; b < NB; b++) {
        const int64_t c = m->boundary2cell_nr[b];
        c2b[c] = (int32_t)b;
        isb[c] = 1;
    }

    // Rays::setup_antipodal_rays (rays.cpp:90-103), within each ray set: the
    // last match wins; antipod holds global ray indices
    ctx->antipod.assign(R, -1);
    for (size_t sidx = 0; sidx + 1 < set_off.size(); sidx++) {
        const int64_t a0 = set_off[sidx], a1 = set_off[sidx + 1];
        for (int64_t r1 = a0; r1 < a1; r1++)
            for (int64_t r2 = a0; r2 < a1; r2++) {
                const double x = m->rays[3 * r1 + 0] + m->rays[3 * r2 + 0];
                const double y = m->rays[3 * r1 + 1] + m->rays[3 * r2 + 1];
                const double z = m->rays[3 * r1 + 2] + m->rays[3 * r2 + 2];
                if (((x * x + y * y) + z * z) < 1.0E-9) ctx->antipod[r1] = r2;
            }
        for (int64_t r = a0; r < a0 + (a1 - a0) / 2; r++)
            if (ctx->antipod[r] < 0) return ctx->fail(MGR_ERR_ARGUMENT, "a ray has no antipode");
    }

    // collision partners of every species, species after species: offsets into
    // the concatenated tables
    std::vector<int32_t> cpn(ctx->ncolpar + 1), cpc(ctx->ncolpar + 1), cpt(ctx->ncolpar + 1),
        cpk(ctx->ncolpar + 1), cptoff(ctx->ncolpar + 1), cpco(ctx->ncolpar + 1), cpop(ctx->ncolpar + 1);
    std::vector<int64_t> icol_all, jcol_all;
    std::vector<double> tmp_all, Cd_all, Ce_all;
    int64_t koff = 0, toff = 0, coff = 0;
    for (int64_t l = 0; l < S; l++) {
        const mgr_species_view &v = spv[l];
        int64_t kl = 0, tl = 0, cl = 0;
        for (int64_t cc = 0; cc < v.ncolpar; cc++) {
            const int64_t c = ctx->sp_cp0[l] + cc;
            if (v.colpar_num[cc] < 0 || v.colpar_num[cc] >= m->nspecs)
                return ctx->fail(MGR_ERR_ARGUMENT, "collision partner species out of range");
            if (v.colpar_ntmp[cc] < 1) return ctx->fail(MGR_ERR_ARGUMENT, "collision partner without temperatures");
            cpn[c] = (int32_t)v.colpar_num[cc]; cpc[c] = (int32_t)v.colpar_ncol[cc];
            cpt[c] = (int32_t)v.colpar_ntmp[cc];
            cpk[c] = (int32_t)koff; cptoff[c] = (int32_t)toff; cpco[c] = (int32_t)coff;
            const char op = v.colpar_orth_or_para[cc];
            cpop[c] = (op == 'o') ? 1 : (op == 'p') ? 2 : 0;
            for (int64_t k = 0; k < v.colpar_ncol[cc]; k++) {
                if (v.icol[kl + k] < 0 || v.icol[kl + k] >= v.nlev || v.jcol[kl + k] < 0 || v.jcol[kl + k] >= v.nlev)
                    return ctx->fail(MGR_ERR_ARGUMENT, "collisional transition level out of range");
                icol_all.push_back(v.icol[kl + k]);
                jcol_all.push_back(v.jcol[kl + k]);
            }
            tmp_all.insert(tmp_all.end(), v.tmp + tl, v.tmp + tl + v.colpar_ntmp[cc]);
            Cd_all.insert(Cd_all.end(), v.Cd + cl, v.Cd + cl + v.colpar_ncol[cc] * v.colpar_ntmp[cc]);
            Ce_all.insert(Ce_all.end(), v.Ce + cl, v.Ce + cl + v.colpar_ncol[cc] * v.colpar_ntmp[cc]);
            kl += v.colpar_ncol[cc]; tl += v.colpar_ntmp[cc]; cl += v.colpar_ncol[cc] * v.colpar_ntmp[cc];
            koff += v.colpar_ncol[cc]; toff += v.colpar_ntmp[cc];
            coff += v.colpar_ncol[cc] * v.colpar_ntmp[cc];
        }
    }
    ctx->ncol_tot = koff; ctx->ntmp_tot = toff; ctx->ctab_tot = coff;

    if ((rc = upload(ctx, &ctx->pos, m->position, N * 3))) return rc;
    if ((rc = upload(ctx, &ctx->vel, m->velocity, N * 3))) return rc;
    if ((rc = upload(ctx, &ctx->nb_off, off.data(), N + 1))) return rc;
    if ((rc = upload_cast<int32_t>(ctx, &ctx->nb, m->neighbors, M))) return rc;
    {
        // the tracer's neighbour records: index and position of every entry
        std::vector<NbRec> rec((size_t)std::max<int64_t>(M, 1));
        for (int64_t q = 0; q < M; q++) {
            const int64_t c = m->neighbors[q];
            rec[q].x = m->position[3 * c + 0];
            rec[q].y = m->position[3 * c + 1];
            rec[q].z = m->position[3 * c + 2];
            rec[q].c = c;
        }
        MGR_HIP(hipMalloc((void **)&ctx->nb_rec, rec.size() * sizeof(NbRec)));
        MGR_HIP(hipMemcpy(ctx->nb_rec, rec.data(), rec.size() * sizeof(NbRec), hipMemcpyHostToDevice));
        // the tracer's candidate masks: list lengths, and a margin of 1e-9 of
        // the largest |p - o|_1 |n|_inf over cells and rays
        ctx->nb_kmax = 0;
        for (int64_t c = 0; c < N; c++) ctx->nb_kmax = std::max<int64_t>(ctx->nb_kmax, off[c + 1] - off[c]);
        double ext = 0.0, nmax = 0.0;
        for (int64_t c = 0; c < N; c++)
            ext = std::max(ext, std::fabs(m->position[3 * c]) + std::fabs(m->position[3 * c + 1]) +
                                    std::fabs(m->position[3 * c + 2]));
        for (int64_t i = 0; i < 3 * R; i++) nmax = std::max(nmax, std::fabs(m->rays[i]));
        ctx->fwd_tau = 1e-9 * (2.0 * ext) * nmax;
    }
    if ((rc = upload(ctx, &ctx->is_boundary, isb.data(), N))) return rc;
    if ((rc = upload(ctx, &ctx->cell2bdy, c2b.data(), N))) return rc;
    if ((rc = upload_cast<int32_t>(ctx, &ctx->bdy2cell, m->boundary2cell_nr, NB))) return rc;
    if ((rc = upload(ctx, &ctx->rays, m->rays, R * 3))) return rc;
    if ((rc = upload(ctx, &ctx->ray_w, m->weights, R))) return rc;
    if ((rc = upload_cast<int32_t>(ctx, &ctx->d_antipod, ctx->antipod.data(), R))) return rc;
    if ((rc = upload(ctx, &ctx->d_cell_model, cell_model.data(), N))) return rc;
    if ((rc = upload(ctx, &ctx->d_cell_qbase, ctx->h_cell_qbase.data(), N))) return rc;
    if ((rc = upload(ctx, &ctx->d_m_R, ctx->m_R.data(), Mm))) return rc;
    if ((rc = upload(ctx, &ctx->d_m_rbase, ctx->m_rbase.data(), Mm))) return rc;
    if ((rc = upload(ctx, &ctx->d_chunk_cell, chunk_cell.data(), ctx->nchunks + 1))) return rc;
    if ((rc = upload(ctx, &ctx->d_m_chunk_off, ctx->m_chunk_off.data(), Mm + 1))) return rc;
    if ((rc = upload(ctx, &ctx->d_m_cell_off, ctx->m_cell_off.data(), Mm + 1))) return rc;
    {
        std::vector<int32_t> ones(Mm, 1);
        if ((rc = upload(ctx, &ctx->d_m_active, ones.data(), Mm))) return rc;
    }
    if ((rc = dalloc(ctx, &ctx->d_mconv, Mm * 4))) return rc;
    if ((rc = dalloc(ctx, &ctx->d_mng, Mm * 4))) return rc;
    MGR_HIP(hipHostMalloc((void **)&ctx->h_m_active, Mm * sizeof(int32_t), hipHostMallocDefault));
    MGR_HIP(hipHostMalloc((void **)&ctx->h_mconv, Mm * 4 * sizeof(double), hipHostMallocDefault));
    for (int64_t k = 0; k < Mm; k++) ctx->h_m_active[k] = 1;
    ctx->m_iterations.assign(Mm, 0);
    ctx->m_last_conv.assign(Mm, mgr_conv{0.0, 0.0, 0.0});
    if ((rc = upload(ctx, &ctx->temperature, m->temperature, N))) return rc;
    if ((rc = upload(ctx, &ctx->vturb2, m->vturb2, N))) return rc;
    if ((rc = upload(ctx, &ctx->abundance, m->abundance, N * m->nspecs))) return rc;
    if ((rc = upload(ctx, &ctx->tables, &tb, 1))) return rc;
    if ((rc = upload(ctx, &ctx->A, A_all.data(), L))) return rc;
    if ((rc = upload(ctx, &ctx->Bs, Bs_all.data(), L))) return rc;
    if ((rc = upload(ctx, &ctx->Ba, Ba_all.data(), L))) return rc;
    if ((rc = upload(ctx, &ctx->frequency, freq_all.data(), L))) return rc;
    if ((rc = upload_cast<int32_t>(ctx, &ctx->irad, irad_all.data(), L))) return rc;
    if ((rc = upload_cast<int32_t>(ctx, &ctx->jrad, jrad_all.data(), L))) return rc;
    if ((rc = upload(ctx, &ctx->energy, energy_all.data(), V))) return rc;
    if ((rc = upload(ctx, &ctx->lweight, weight_all.data(), V))) return rc;
    if ((rc = upload(ctx, &ctx->roots, roots_all.data(), S * Q))) return rc;
    ctx->root_max = 0.0;
    for (int64_t z = 0; z < S * Q; z++) ctx->root_max = std::max(ctx->root_max, std::fabs(roots_all[z]));
    if ((rc = upload(ctx, &ctx->qweights, qw_all.data(), S * Q))) return rc;
    if ((rc = upload(ctx, &ctx->cp_num, cpn.data(), ctx->ncolpar))) return rc;
    ctx->h_cp_num.assign(cpn.begin(), cpn.begin() + ctx->ncolpar);
    if ((rc = upload(ctx, &ctx->cp_ncol, cpc.data(), ctx->ncolpar))) return rc;
    if ((rc = upload(ctx, &ctx->cp_ntmp, cpt.data(), ctx->ncolpar))) return rc;
    if ((rc = upload(ctx, &ctx->cp_koff, cpk.data(), ctx->ncolpar))) return rc;
    if ((rc = upload(ctx, &ctx->cp_toff, cptoff.data(), ctx->ncolpar))) return rc;
    if ((rc = upload(ctx, &ctx->cp_coff, cpco.data(), ctx->ncolpar))) return rc;
    if ((rc = upload(ctx, &ctx->cp_op, cpop.data(), ctx->ncolpar))) return rc;
    if ((rc = upload_cast<int32_t>(ctx, &ctx->icol, icol_all.data(), koff))) return rc;
    if ((rc = upload_cast<int32_t>(ctx, &ctx->jcol, jcol_all.data(), koff))) return rc;
    if ((rc = upload(ctx, &ctx->tmp, tmp_all.data(), toff))) return rc;
    if ((rc = upload(ctx, &ctx->Cd, Cd_all.data(), coff))) return rc;
    if ((rc = upload(ctx, &ctx->Ce, Ce_all.data(), coff))) return rc;

    // derived and state arrays
    const int64_t F = ctx->F;
    if ((rc = dalloc(ctx, &ctx->pw, S * N))) return rc;
    if ((rc = dalloc(ctx, &ctx->pw_cmax, N))) return rc;
    if ((rc = dalloc(ctx, &ctx->dshift_max, N))) return rc;
    if ((rc = dalloc(ctx, &ctx->W11, N * L))) return rc;
    if ((rc = dalloc(ctx, &ctx->IW64, N * L))) return rc;
    if ((rc = dalloc(ctx, &ctx->IW32, N * L))) return rc;
    if ((rc = dalloc(ctx, &ctx->row32, N * L))) return rc;
    if ((rc = dalloc(ctx, &ctx->row64, N * L))) return rc;
    if ((rc = dalloc(ctx, &ctx->nu, N * F))) return rc;
    if ((rc = dalloc(ctx, &ctx->nr_line, N * L * Q))) return rc;
    if ((rc = dalloc(ctx, &ctx->I_bdy, NB * F))) return rc;
    if ((rc = dalloc(ctx, &ctx->pop, N * V))) return rc;
    if ((rc = dalloc(ctx, &ctx->prev1, N * V))) return rc;
    if ((rc = dalloc(ctx, &ctx->prev2, N * V))) return rc;
    if ((rc = dalloc(ctx, &ctx->prev3, N * V))) return rc;
    if ((rc = dalloc(ctx, &ctx->pop_tot, S * N))) return rc;
    if ((rc = dalloc(ctx, &ctx->emis, N * L))) return rc;
    if ((rc = dalloc(ctx, &ctx->opac, N * L))) return rc;
    if ((rc = dalloc(ctx, &ctx->J, N * F))) return rc;
    if ((rc = dalloc(ctx, &ctx->Lam, N * F))) return rc;
    if ((rc = dalloc(ctx, &ctx->Ls, N * L))) return rc;
    if ((rc = dalloc(ctx, &ctx->Lpart, N * ((F + WAVE - 1) / WAVE) * L))) return rc;
    if ((rc = dalloc(ctx, &ctx->Jlin, N * L))) return rc;
    if ((rc = dalloc(ctx, &ctx->Jeff, N * L))) return rc;
    if ((rc = dalloc(ctx, &ctx->Jdif, N * L))) return rc;
    if (ctx->nband > 0) {
        const int64_t S = ctx->band_S;
        if ((rc = dalloc(ctx, &ctx->band_cell, N * S))) return rc;
        if ((rc = dalloc(ctx, &ctx->band_part, N * ((F + WAVE - 1) / WAVE) * S * L))) return rc;
        if ((rc = dalloc(ctx, &ctx->band_Ls, N * S * L))) return rc;
        if ((rc = dalloc(ctx, &ctx->band_lu, N * V * V))) return rc;
        if ((rc = dalloc(ctx, &ctx->band_piv, N * V))) return rc;
        if ((rc = dalloc(ctx, &ctx->band_x, N * V))) return rc;
        if ((rc = dalloc(ctx, &ctx->band_flags, BAND_SWEEPS + 1))) return rc;
        if ((rc = dalloc(ctx, &ctx->band_change, BAND_SWEEPS))) return rc;
        MGR_HIP(hipMemset(ctx->band_Ls, 0, N * S * L * sizeof(double)));
    }
    if ((rc = dalloc(ctx, &ctx->err_flag, 1))) return rc;
    if ((rc = dalloc(ctx, &ctx->partials, (size_t)ctx->nchunks * 8))) return rc;
    if ((rc = dalloc(ctx, &ctx->reduced, 4 * (MAX_SPECIES + 1)))) return rc;
    if ((rc = dalloc(ctx, &ctx->shift_bits, 2))) return rc;
    MGR_HIP(hipHostMalloc((void **)&ctx->h_reduced, 8 * sizeof(double), hipHostMallocDefault));
    MGR_HIP(hipMemset(ctx->err_flag, 0, sizeof(int32_t)));
    MGR_HIP(hipMemset(ctx->prev1, 0, N * V * sizeof(double)));
    MGR_HIP(hipMemset(ctx->prev2, 0, N * V * sizeof(double)));
    MGR_HIP(hipMemset(ctx->prev3, 0, N * V * sizeof(double)));
    MGR_HIP(hipMemset(ctx->Ls, 0, N * L * sizeof(double)));
    MGR_HIP(hipMemset(ctx->J, 0, N * F * sizeof(double)));

    for (int64_t l = 0; l < S; l++) {
        hipLaunchKernelGGL(k_pop_tot, dim3(grid_for(N, 256)), dim3(256), 0, ctx->stream, N,
                           (int)ctx->nspecs, (int)ctx->sp_num[l], ctx->abundance, ctx->pop_tot + l * N);
        MGR_LAUNCH_CHECK();
    }
    if ((rc = launch_cell_tables(ctx))) return rc;
    MGR_HIP(hipStreamSynchronize(ctx->stream));
    {
        std::vector<double> pwh(S * N);
        MGR_HIP(hipMemcpy(pwh.data(), ctx->pw, S * N * sizeof(double), hipMemcpyDeviceToHost));
        ctx->pw_max = *std::max_element(pwh.begin(), pwh.end());
    }
    return MGR_OK;
}

int free_model(mgr_ctx *ctx)
{
    dfree(ctx->pos); dfree(ctx->vel); dfree(ctx->nb_off); dfree(ctx->nb); dfree(ctx->is_boundary);
    dfree(ctx->nb_rec); dfree(ctx->fwd_mask);
    dfree(ctx->cell2bdy); dfree(ctx->bdy2cell); dfree(ctx->rays); dfree(ctx->ray_w);
    dfree(ctx->d_antipod); dfree(ctx->temperature); dfree(ctx->vturb2); dfree(ctx->abundance);
    dfree(ctx->tables); dfree(ctx->A); dfree(ctx->Bs); dfree(ctx->Ba); dfree(ctx->frequency);
    dfree(ctx->irad); dfree(ctx->jrad); dfree(ctx->energy); dfree(ctx->lweight); dfree(ctx->roots);
    dfree(ctx->qweights); dfree(ctx->cp_num); dfree(ctx->cp_ncol); dfree(ctx->cp_ntmp);
    dfree(ctx->cp_koff); dfree(ctx->cp_toff); dfree(ctx->cp_coff); dfree(ctx->cp_op);
    dfree(ctx->icol); dfree(ctx->jcol); dfree(ctx->tmp); dfree(ctx->Cd); dfree(ctx->Ce);
    dfree(ctx->pw); dfree(ctx->pw_cmax); dfree(ctx->dshift_max); dfree(ctx->W11); dfree(ctx->IW64); dfree(ctx->IW32);
    dfree(ctx->row32); dfree(ctx->row64); dfree(ctx->nu);
    dfree(ctx->nr_line); dfree(ctx->f_appears); dfree(ctx->f_k); dfree(ctx->f_z);
    dfree(ctx->d_lam_f_off); dfree(ctx->d_lam_f); dfree(ctx->I_bdy); dfree(ctx->pop);
    dfree(ctx->prev1); dfree(ctx->prev2); dfree(ctx->prev3); dfree(ctx->pop_tot); dfree(ctx->emis);
    dfree(ctx->opac); dfree(ctx->J); dfree(ctx->Lam); dfree(ctx->Ls); dfree(ctx->Lpart); dfree(ctx->Jlin);
    dfree(ctx->Jeff); dfree(ctx->Jdif); dfree(ctx->err_flag); dfree(ctx->seg_off); dfree(ctx->nseg);
    dfree(ctx->npts); dfree(ctx->seg_cell); dfree(ctx->seg_shift); dfree(ctx->seg_dZ);
    dfree(ctx->seg_f32); dfree(ctx->seg_hdz32); dfree(ctx->far_cell); dfree(ctx->far_shift);
    dfree(ctx->step_mask); dfree(ctx->step_mask64);
    dfree(ctx->far_I32);
    dfree(ctx->far_I64);
    dfree(ctx->shift_bits); dfree(ctx->partials); dfree(ctx->reduced);
    dfree(ctx->d_cell_model); dfree(ctx->d_cell_qbase); dfree(ctx->d_m_R); dfree(ctx->d_m_rbase);
    dfree(ctx->d_m_active); dfree(ctx->d_chunk_cell); dfree(ctx->d_m_chunk_off); dfree(ctx->d_mconv);
    dfree(ctx->d_m_cell_off);
    dfree(ctx->d_mng);
    dfree(ctx->u_store);
    dfree(ctx->band_cell); dfree(ctx->band_part); dfree(ctx->band_Ls); dfree(ctx->band_rev_off);
    dfree(ctx->band_rev); dfree(ctx->band_lu); dfree(ctx->band_piv); dfree(ctx->band_x); dfree(ctx->band_flags);
    dfree(ctx->band_change);
    free_image(ctx);
    if (ctx->h_reduced) (void)hipHostFree(ctx->h_reduced);
    if (ctx->h_m_active) (void)hipHostFree(ctx->h_m_active);
    if (ctx->h_mconv) (void)hipHostFree(ctx->h_mconv);
    ctx->h_reduced = nullptr;
    ctx->h_m_active = nullptr;
    ctx->h_mconv = nullptr;
    if (ctx->graph_normal) (void)hipGraphExecDestroy(ctx->graph_normal);
    if (ctx->graph_ng) (void)hipGraphExecDestroy(ctx->graph_ng);
    ctx->graph_normal = ctx->graph_ng = nullptr;
    return MGR_OK;
}

}  // namespace mgr

// ---- C-ABI -------------------------------------------------------------------

extern "C" {

int mgr_device_count(int *count)
{
    if (!count) return MGR_ERR_ARGUMENT;
    *count = 0;
    return hipGetDeviceCount(count) == hipSuccess ? MGR_OK : MGR_ERR_DEVICE;
}

int mgr_create_batch(const mgr_model_view *models, int64_t nmodels, int device, mgr_ctx **out)
{
    if (!out) return MGR_ERR_ARGUMENT;
    *out = nullptr;
    if (!models || nmodels < 1) { g_create_error = "null or empty model list"; return MGR_ERR_ARGUMENT; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        g_create_error = "no HIP device available";
        return MGR_ERR_DEVICE;
    }
    if (device < 0 || device >= ndev) { g_create_error = "device index out of range"; return MGR_ERR_ARGUMENT; }
    mgr_ctx *ctx = new mgr_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        g_create_error = "could not set device / create stream";
        delete ctx;
        return MGR_ERR_DEVICE;
    }
    const int rc = upload_models(ctx, models, nmodels);
    if (rc != MGR_OK) {
        g_create_error = ctx->err;
        free_model(ctx);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return rc;
    }
    *out = ctx;
    return MGR_OK;
}

int mgr_create(const mgr_model_view *model, int device, mgr_ctx **out)
{
    if (!model) {
        if (out) *out = nullptr;
        g_create_error = "null model";
        return MGR_ERR_ARGUMENT;
    }
    return mgr_create_batch(model, 1, device, out);
}

int mgr_batch_info(mgr_ctx *ctx, int64_t *nmodels, int64_t *cell_offsets)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    if (nmodels) *nmodels = ctx->M;
    if (cell_offsets) std::memcpy(cell_offsets, ctx->m_cell_off.data(), (ctx->M + 1) * sizeof(int64_t));
    return MGR_OK;
}

int mgr_batch_convergence(mgr_ctx *ctx, int64_t *iterations, mgr_conv *conv)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    if (iterations) std::memcpy(iterations, ctx->m_iterations.data(), ctx->M * sizeof(int64_t));
    if (conv) std::memcpy(conv, ctx->m_last_conv.data(), ctx->M * sizeof(mgr_conv));
    return MGR_OK;
}

static int copy_out(mgr_ctx *ctx, void *dst, const void *src, size_t bytes);

int mgr_species_info(mgr_ctx *ctx, int64_t *nlspecs, int64_t *nlev, int64_t *nrad)
{
    if (!ctx || !nlspecs) return MGR_ERR_ARGUMENT;
    *nlspecs = ctx->S;
    for (int64_t l = 0; l < ctx->S; l++) {
        if (nlev) nlev[l] = ctx->sp_V[l];
        if (nrad) nrad[l] = ctx->sp_L[l];
    }
    return MGR_OK;
}

int mgr_species_convergence(mgr_ctx *ctx, int64_t l, mgr_conv *conv)
{
    if (!ctx || !conv) return MGR_ERR_ARGUMENT;
    if (l < 0 || l >= ctx->S) return ctx->fail(MGR_ERR_ARGUMENT, "species index out of range");
    double w[3];
    const int rc = copy_out(ctx, w, ctx->reduced + (ctx->S > 1 ? 4 * (l + 1) : 0), sizeof(w));
    if (rc) return rc;
    conv->fraction_not_converged = w[0];
    conv->relative_change_mean = w[1];
    conv->relative_change_max = w[2];
    return MGR_OK;
}

void mgr_destroy(mgr_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    free_model(ctx);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *mgr_last_error(const mgr_ctx *ctx)
{
    if (!ctx) return g_create_error.c_str();
    return ctx->err.c_str();
}

static int spectral_discretisation(mgr_ctx *ctx, bool image, double width)
{
    MGR_HIP(hipSetDevice(ctx->device));
    const int64_t N = ctx->N, F = ctx->F, L = ctx->L, Q = ctx->Q;
    if (F * (int64_t)sizeof(double) <= 48 * 1024)
        hipLaunchKernelGGL(k_spectral_cell, dim3((unsigned)N), dim3(256), F * sizeof(double), ctx->stream, N,
                           ctx->tables, image ? nullptr : ctx->pw, width, ctx->roots, ctx->nu, ctx->nr_line);
    else
        hipLaunchKernelGGL(k_spectral, dim3(grid_for(N * F, 256)), dim3(256), 0, ctx->stream, N,
                           ctx->tables, image ? nullptr : ctx->pw, width, ctx->roots, ctx->nu, ctx->nr_line);
    MGR_LAUNCH_CHECK();
    // appears_in_line_integral / corresponding_{l,k,z}, indexed by the unsorted
    // position (sim_spectrum.tpp:153-184), and per line the sorted frequency
    // indices f whose back-map names it (used for Lambda*).  Here f_k holds
    // the line g = loff(l) + k and f_z the index l Q + z of the root's weight
    // in the species' quadratures held one after the other.
    std::vector<int32_t> appears(F, 0), fk(F, (int32_t)F), fz(F, (int32_t)F);
    for (int64_t k = 0; k < L; k++)
        for (int64_t z = 0; z < Q; z++) {
            appears[k * Q + z] = 1;
            fk[k * Q + z] = (int32_t)k;
            fz[k * Q + z] = (int32_t)(ctx->h_tables.line_spec[k] * Q + z);
        }
    ctx->lam_f_off.assign(L + 1, 0);
    ctx->lam_f.clear();
    for (int64_t k = 0; k < L; k++) {
        for (int64_t f = 0; f < F; f++)
            if (appears[f] && fk[f] == k) ctx->lam_f.push_back(f);
        ctx->lam_f_off[k + 1] = (int64_t)ctx->lam_f.size();
    }
    int rc;
    std::vector<int32_t> off32(ctx->lam_f_off.begin(), ctx->lam_f_off.end()), f32(ctx->lam_f.begin(), ctx->lam_f.end());
    mgr_ctx *c = ctx;
    if (!c->f_appears) {
        if ((rc = upload(c, &c->f_appears, appears.data(), F))) return rc;
        if ((rc = upload(c, &c->f_k, fk.data(), F))) return rc;
        if ((rc = upload(c, &c->f_z, fz.data(), F))) return rc;
        if ((rc = upload(c, &c->d_lam_f_off, off32.data(), L + 1))) return rc;
        if ((rc = upload(c, &c->d_lam_f, f32.data(), f32.size()))) return rc;
    }
    MGR_HIP(hipStreamSynchronize(ctx->stream));
    ctx->spectral_set = true;
    ctx->spectral_image = image;
    ctx->masks_valid = false;   // the out-of-window bits depend on the frequencies
    ctx->masks64_valid = false;
    ctx->far_invalidate();
    return MGR_OK;
}

int mgr_compute_spectral_discretisation(mgr_ctx *ctx)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    return spectral_discretisation(ctx, false, 0.0);
}

int mgr_compute_spectral_discretisation_image(mgr_ctx *ctx, double width)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    if (!(width > 0.0)) return ctx->fail(MGR_ERR_ARGUMENT, "image width must be positive");
    return spectral_discretisation(ctx, true, width);
}

int mgr_compute_boundary_intensities(mgr_ctx *ctx, const double *temperatures)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    if (!ctx->spectral_set) return ctx->fail(MGR_ERR_STATE, "compute_spectral_discretisation first");
    MGR_HIP(hipSetDevice(ctx->device));
    double *dT = nullptr;
    if (temperatures) {
        int rc = upload(ctx, &dT, temperatures, ctx->NB);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_boundary_intensities, dim3(grid_for(ctx->NB * ctx->F, 256)), dim3(256), 0,
                       ctx->stream, ctx->NB, (int)ctx->F, ctx->bdy2cell, dT, T_CMB, ctx->nu, ctx->I_bdy);
    hipError_t le = hipGetLastError();
    hipError_t se = hipStreamSynchronize(ctx->stream);
    if (dT) (void)hipFree(dT);
    if (le != hipSuccess || se != hipSuccess)
        return ctx->fail(MGR_ERR_DEVICE, "boundary intensities kernel failed");
    ctx->bdy_set = true;
    ctx->far_invalidate();
    return MGR_OK;
}

int mgr_compute_LTE_level_populations(mgr_ctx *ctx)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    MGR_HIP(hipSetDevice(ctx->device));
    for (int64_t l = 0; l < ctx->S; l++) {   // every species (sim_lines.tpp:10-16)
        hipLaunchKernelGGL(k_lte, dim3(grid_for(ctx->N, 256)), dim3(256), 0, ctx->stream, ctx->N,
                           (int)ctx->sp_V[l], (int)ctx->nspecs, (int)ctx->sp_num[l], ctx->abundance,
                           ctx->temperature, ctx->lweight + ctx->sp_voff[l], ctx->energy + ctx->sp_voff[l],
                           ctx->pop + ctx->N * ctx->sp_voff[l], ctx->pop_tot + ctx->N * l);
        MGR_LAUNCH_CHECK();
    }
    int rc = launch_emissivity_opacity(ctx);
    if (rc) return rc;
    MGR_HIP(hipStreamSynchronize(ctx->stream));
    ctx->pops_set = true;
    return MGR_OK;
}

static int copy_out(mgr_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (!dst) return MGR_OK;
    MGR_HIP(hipSetDevice(ctx->device));
    MGR_HIP(hipStreamSynchronize(ctx->stream));
    MGR_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return MGR_OK;
}

static int copy_in(mgr_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (!src) return MGR_OK;
    MGR_HIP(hipSetDevice(ctx->device));
    MGR_HIP(hipStreamSynchronize(ctx->stream));
    MGR_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return MGR_OK;
}

int mgr_set_populations(mgr_ctx *ctx, const double *population)
{
    if (!ctx || !population) return MGR_ERR_ARGUMENT;
    int rc = copy_in(ctx, ctx->pop, population, ctx->N * ctx->V * sizeof(double));
    if (rc) return rc;
    if ((rc = launch_emissivity_opacity(ctx))) return rc;
    MGR_HIP(hipStreamSynchronize(ctx->stream));
    ctx->pops_set = true;
    return MGR_OK;
}

int mgr_get_populations(mgr_ctx *ctx, double *population)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    return copy_out(ctx, population, ctx->pop, ctx->N * ctx->V * sizeof(double));
}

int mgr_get_population_history(mgr_ctx *ctx, double *p1, double *p2, double *p3)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    const size_t n = ctx->N * ctx->V * sizeof(double);
    int rc;
    if ((rc = copy_out(ctx, p1, ctx->prev1, n))) return rc;
    if ((rc = copy_out(ctx, p2, ctx->prev2, n))) return rc;
    return copy_out(ctx, p3, ctx->prev3, n);
}

int mgr_set_population_history(mgr_ctx *ctx, const double *p1, const double *p2, const double *p3)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    const size_t n = ctx->N * ctx->V * sizeof(double);
    int rc;
    if ((rc = copy_in(ctx, ctx->prev1, p1, n))) return rc;
    if ((rc = copy_in(ctx, ctx->prev2, p2, n))) return rc;
    return copy_in(ctx, ctx->prev3, p3, n);
}

int mgr_get_J(mgr_ctx *ctx, double *J)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    return copy_out(ctx, J, ctx->J, ctx->N * ctx->F * sizeof(double));
}

int mgr_get_lambda(mgr_ctx *ctx, double *Ls)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    return copy_out(ctx, Ls, ctx->Ls, ctx->N * ctx->L * sizeof(double));
}

int mgr_lambda_band_slots(mgr_ctx *ctx, int64_t *slots)
{
    if (!ctx || !slots) return MGR_ERR_ARGUMENT;
    *slots = ctx->band_S;
    return MGR_OK;
}

int mgr_band_sweeps(mgr_ctx *ctx, int *sweeps, double *changes)
{
    if (!ctx || !sweeps) return MGR_ERR_ARGUMENT;
    if (ctx->nband <= 0) return ctx->fail(MGR_ERR_STATE, "the model has no Lambda* bands (n_off_diag < 2)");
    MGR_HIP(hipSetDevice(ctx->device));
    MGR_HIP(hipStreamSynchronize(ctx->stream));
    unsigned long long bits[BAND_SWEEPS];
    MGR_HIP(hipMemcpy(bits, ctx->band_change, sizeof(bits), hipMemcpyDeviceToHost));
    int n = 0;
    for (int t = 0; t < BAND_SWEEPS; t++) {
        double v;
        std::memcpy(&v, &bits[t], sizeof(v));
        if (changes) changes[t] = v;
        if (bits[t]) n = t + 1;
    }
    *sweeps = n;
    return MGR_OK;
}

int mgr_get_lambda_band(mgr_ctx *ctx, int32_t *cells, double *Ls)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    if (ctx->nband <= 0) return ctx->fail(MGR_ERR_STATE, "the model has no Lambda* bands (n_off_diag < 2)");
    if (!ctx->traced) return ctx->fail(MGR_ERR_STATE, "mgr_trace first");
    int rc;
    if ((rc = copy_out(ctx, cells, ctx->band_cell, ctx->N * ctx->band_S * sizeof(int32_t)))) return rc;
    return copy_out(ctx, Ls, ctx->band_Ls, ctx->N * ctx->band_S * ctx->L * sizeof(double));
}

int mgr_get_Jeff(mgr_ctx *ctx, double *Jlin, double *Jeff, double *Jdif)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    const size_t n = ctx->N * ctx->L * sizeof(double);
    int rc;
    if ((rc = copy_out(ctx, Jlin, ctx->Jlin, n))) return rc;
    if ((rc = copy_out(ctx, Jeff, ctx->Jeff, n))) return rc;
    return copy_out(ctx, Jdif, ctx->Jdif, n);
}

int mgr_get_frequencies(mgr_ctx *ctx, double *nu, int64_t *nr_line)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    if (!ctx->spectral_set) return ctx->fail(MGR_ERR_STATE, "compute_spectral_discretisation first");
    int rc = copy_out(ctx, nu, ctx->nu, ctx->N * ctx->F * sizeof(double));
    if (rc || !nr_line) return rc;
    const int64_t n = ctx->N * ctx->L * ctx->Q;
    std::vector<int32_t> tmpv(n);
    if ((rc = copy_out(ctx, tmpv.data(), ctx->nr_line, n * sizeof(int32_t)))) return rc;
    for (int64_t i = 0; i < n; i++) nr_line[i] = tmpv[i];
    return MGR_OK;
}

int mgr_get_boundary_intensities(mgr_ctx *ctx, double *I_bdy)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    return copy_out(ctx, I_bdy, ctx->I_bdy, ctx->NB * ctx->F * sizeof(double));
}

int mgr_get_emissivity_opacity(mgr_ctx *ctx, double *emissivity, double *opacity)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    const size_t n = ctx->N * ctx->L * sizeof(double);
    int rc = copy_out(ctx, emissivity, ctx->emis, n);
    if (rc) return rc;
    return copy_out(ctx, opacity, ctx->opac, n);
}

int mgr_last_radiation_field_ms(mgr_ctx *ctx, double *ms)
{
    if (!ctx || !ms) return MGR_ERR_ARGUMENT;
    *ms = ctx->last_rf_ms;
    return MGR_OK;
}

int mgr_last_radiation_field_kernel(mgr_ctx *ctx, char *name, int capacity)
{
    if (!ctx || !name || capacity < 1) return MGR_ERR_ARGUMENT;
    std::snprintf(name, (size_t)capacity, "%s", ctx->last_rf_kernel);
    return MGR_OK;
}

int mgr_trace_stats(mgr_ctx *ctx, int64_t *sum_ndep, int64_t *n_pairs_solved)
{
    if (!ctx) return MGR_ERR_ARGUMENT;
    if (!ctx->traced) return ctx->fail(MGR_ERR_STATE, "mgr_trace first");
    if (sum_ndep) *sum_ndep = ctx->sum_ndep;
    if (n_pairs_solved) *n_pairs_solved = ctx->n_pairs_solved;
    return MGR_OK;
}

}  // extern "C"
