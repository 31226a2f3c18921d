/*
 * oracle.c -- CPU FP64 restatement ("cpu_ref") of Magritte's RayPair
 * radiation-field path and ALI driver.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the timed CPU
 * baseline.  Nothing in the product links or loads it.
 *
 * Built with -O2 -ffp-contract=off (no fused multiply-add), scalar FP64,
 * OpenMP over origin cells exactly as the reference: ray pairs outer, origins
 * strided over threads (src/Simulation/sim_radiation.tpp:130-148).
 *
 * Deviations, all stated where they occur:
 *  - the global Eigen SparseLU<COLAMD> of the rate equations is replaced by one
 *    dense LU with partial pivoting per cell; with n_off_diag = 0 every triplet
 *    couples levels of one cell, so the global matrix is block diagonal and the
 *    solutions agree up to rounding;
 *  - heapsort (src/Functions/heapsort.cpp:77-114) is replaced by a sort on
 *    (value, original index); for distinct values the permutation is the same;
 *  - relative_change_max is computed (the reference leaves it at 0 because its
 *    update is commented out as not thread safe, lineProducingSpecies.tpp:157-161);
 *  - a segment needing more than 10000 interpolation points contributes no
 *    points, as set_data's early return does (geometry.tpp:138-147);
 *  - a ray with no forward neighbour stops and is reported (-1) instead of
 *    reading velocity[-1] (geometry.tpp:68-77).
 *  - with n_off_diag >= 2 the band elements (raypair_lambda.tpp:129-146) are
 *    added only where the solve of the same pair wrote them: the lower band m
 *    reads L_lower[m][n_ar-m], written iff the shorter ray has a point
 *    (n_r >= 1), and the upper band m reads L_upper[m][n_ar+m], written iff
 *    n_r > 2m (raypair_solver.tpp:381-388).  Elsewhere the reference reads what
 *    an earlier pair of the same thread left in its buffers; those elements
 *    are left out;
 *  - with n_off_diag >= 2 the global rate equations (cells coupled through the
 *    band elements) are solved with one dense LU with partial pivoting over
 *    all N V unknowns in place of SparseLU<COLAMD>;
 *  - several line-producing species: each species' rate equations are solved
 *    per cell as above (the reference's global matrix is block diagonal over
 *    species as well); bands are restated for one species only.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- constants: src/Tools/constants.hpp:26-63 ---------------------------- */
#define O_PI              3.14159265358979323846
#define O_FOUR_PI         (4.0 * O_PI)
#define O_INVERSE_SQRT_PI (0.5 * 1.12837916709551257390) /* 0.5*M_2_SQRTPI */
#define O_CC              2.99792458E+8
#define O_HH              6.62607004E-34
#define O_KB              1.38064852E-23
#define O_AMU             1.66053904E-27
#define O_T_CMB           2.7254800
#define O_C_SQUARED       8.98755179E+16
static const double TWO_KB_OVER_AMU_C_SQUARED = 2.0 * O_KB / (O_AMU * O_C_SQUARED);
static const double TWO_HH_OVER_CC_SQUARED    = 2.0 * O_HH / O_C_SQUARED;
static const double HH_OVER_KB                = O_HH / O_KB;
static const double HH_OVER_FOUR_PI           = O_HH / O_FOUR_PI;

/* One line-producing species (Lines/LineProducingSpecies_<l>): its fields,
 * level and line counts, and where its populations (from N*voff, index
 * p*V + i) and its lines (global line loff + k) start. */
typedef struct {
    mgr_species_view v;
    int64_t V, L, voff, loff;
    double  fnc, rcm, rcmax;
} o_spec;

struct orc_state {
    mgr_model_view m;
    int64_t N, R, NB, F, L, V, Q;   /* L, V: lines and levels summed over species */
    int64_t S;                      /* line-producing species */
    o_spec *sp;
    double  *line;        int64_t *line_index;            /* [L] sorted line list */
    int64_t *cell2boundary_nr; unsigned char *boundary;   /* [N] */
    int64_t *antipod;                                     /* [R] */
    int64_t *nb_off;                                      /* [N+1] */
    double  *nu;          int64_t *nr_line;               /* [N*F], [N*L*Q] */
    unsigned char *appears; int64_t *corr_l, *corr_k, *corr_z; /* [F] */
    int      spec_set;
    double  *I_bdy;                                       /* [R/2][NB][F] */
    double  *pop, *prev1, *prev2, *prev3;                 /* [N*V], species after species */
    double  *pop_tot;                                     /* [S*N] */
    double  *emissivity, *opacity;                        /* [N*L] */
    double  *J;                                           /* [N*F] */
    double  *Ls;                                          /* [N*L] element nr = p */
    /* Lambda* elements with nr != p (n_off_diag >= 2), per (p,k) in the order
     * Lambda::add_element first met them: band_n[p*L+k] entries at
     * (p*L+k)*band_cap */
    int64_t  band_cap;
    int64_t *band_n, *band_nr;
    double  *band_Ls;
    double  *Jlin, *Jeff, *Jdif;                          /* [N*L] */
    double  *u;                                           /* [R/2][N*F] with scattering, else NULL */
    double   fnc, rcm, rcmax;
};

int orc_nthreads_max(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static void *xcalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

/* Stable sort of (value, index) pairs; stands in for heapsort
 * (src/Functions/heapsort.cpp:77-114), identical for distinct values. */
typedef struct { double v; int64_t i; } vi_pair;
static int vi_cmp(const void *a, const void *b)
{
    const vi_pair *x = (const vi_pair *)a, *y = (const vi_pair *)b;
    if (x->v < y->v) return -1;
    if (x->v > y->v) return 1;
    return (x->i < y->i) ? -1 : (x->i > y->i);
}

orc_state *orc_create(const mgr_model_view *m)
{
    orc_state *s = (orc_state *)xcalloc(1, sizeof(orc_state));
    s->m  = *m;
    s->N  = m->ncells;
    s->R  = m->nrays;
    s->NB = m->nboundary;
    s->Q  = m->nquads;
    /* species 0 from the model view, species 1 .. from more_species */
    s->S  = m->nlspecs > 1 ? m->nlspecs : 1;
    s->sp = (o_spec *)xcalloc(s->S, sizeof(o_spec));
    {
        mgr_species_view *v0 = &s->sp[0].v;
        v0->num = m->num; v0->inverse_mass = m->inverse_mass; v0->nlev = m->nlev; v0->nrad = m->nrad;
        v0->energy = m->energy; v0->weight = m->weight; v0->irad = m->irad; v0->jrad = m->jrad;
        v0->A = m->A; v0->Bs = m->Bs; v0->Ba = m->Ba; v0->frequency = m->frequency;
        v0->ncolpar = m->ncolpar; v0->colpar_num = m->colpar_num;
        v0->colpar_orth_or_para = m->colpar_orth_or_para; v0->colpar_ncol = m->colpar_ncol;
        v0->colpar_ntmp = m->colpar_ntmp; v0->icol = m->icol; v0->jcol = m->jcol; v0->tmp = m->tmp;
        v0->Cd = m->Cd; v0->Ce = m->Ce; v0->roots = m->roots; v0->qweights = m->qweights;
        for (int64_t l = 1; l < s->S; l++) s->sp[l].v = m->more_species[l-1];
        s->V = 0; s->L = 0;
        for (int64_t l = 0; l < s->S; l++) {
            s->sp[l].V = s->sp[l].v.nlev; s->sp[l].L = s->sp[l].v.nrad;
            s->sp[l].voff = s->V; s->sp[l].loff = s->L;
            s->V += s->sp[l].V; s->L += s->sp[l].L;
        }
    }
    s->F  = s->L * s->Q;   /* nfreqs = nlines * nquads */

    /* Lines::setup (src/Model/Lines/lines.cpp:43-87): the lines of every
     * species, species after species, sorted */
    s->line = (double *)xcalloc(s->L, sizeof(double));
    s->line_index = (int64_t *)xcalloc(s->L, sizeof(int64_t));
    {
        vi_pair *tmp = (vi_pair *)xcalloc(s->L, sizeof(vi_pair));
        for (int64_t l = 0; l < s->S; l++)
            for (int64_t k = 0; k < s->sp[l].L; k++) {
                tmp[s->sp[l].loff + k].v = s->sp[l].v.frequency[k];
                tmp[s->sp[l].loff + k].i = s->sp[l].loff + k;
            }
        qsort(tmp, (size_t)s->L, sizeof(vi_pair), vi_cmp);
        for (int64_t k = 0; k < s->L; k++) { s->line[k] = tmp[k].v; s->line_index[k] = tmp[k].i; }
        free(tmp);
    }

    /* Boundary::read (src/Model/Geometry/Boundary/boundary.cpp:19-51) */
    s->cell2boundary_nr = (int64_t *)xcalloc(s->N, sizeof(int64_t));
    s->boundary = (unsigned char *)xcalloc(s->N, 1);
    for (int64_t p = 0; p < s->N; p++) { s->cell2boundary_nr[p] = s->N; s->boundary[p] = 0; }
    for (int64_t b = 0; b < s->NB; b++) {
        const int64_t c = m->boundary2cell_nr[b];
        s->cell2boundary_nr[c] = b;
        s->boundary[c] = 1;
    }

    /* Rays::setup_antipodal_rays (src/Model/Geometry/Rays/rays.cpp:90-103):
     * the last r2 with |r1+r2|^2 < 1e-9 wins. */
    s->antipod = (int64_t *)xcalloc(s->R, sizeof(int64_t));
    for (int64_t r1 = 0; r1 < s->R; r1++) {
        for (int64_t r2 = 0; r2 < s->R; r2++) {
            const double x = m->rays[3*r1+0] + m->rays[3*r2+0];
            const double y = m->rays[3*r1+1] + m->rays[3*r2+1];
            const double z = m->rays[3*r1+2] + m->rays[3*r2+2];
            if (((x*x + y*y) + z*z) < 1.0E-9) s->antipod[r1] = r2;
        }
    }

    s->nb_off = (int64_t *)xcalloc(s->N + 1, sizeof(int64_t));
    for (int64_t p = 0; p < s->N; p++) s->nb_off[p+1] = s->nb_off[p] + m->n_neighbors[p];

    s->nu       = (double *)xcalloc(s->N * s->F, sizeof(double));
    s->nr_line  = (int64_t *)xcalloc(s->N * s->L * s->Q, sizeof(int64_t));
    s->appears  = (unsigned char *)xcalloc(s->F, 1);
    s->corr_l   = (int64_t *)xcalloc(s->F, sizeof(int64_t));
    s->corr_k   = (int64_t *)xcalloc(s->F, sizeof(int64_t));
    s->corr_z   = (int64_t *)xcalloc(s->F, sizeof(int64_t));
    s->I_bdy    = (double *)xcalloc((s->R / 2) * s->NB * s->F, sizeof(double));
    s->pop      = (double *)xcalloc(s->N * s->V, sizeof(double));
    s->prev1    = (double *)xcalloc(s->N * s->V, sizeof(double));
    s->prev2    = (double *)xcalloc(s->N * s->V, sizeof(double));
    s->prev3    = (double *)xcalloc(s->N * s->V, sizeof(double));
    s->pop_tot  = (double *)xcalloc(s->S * s->N, sizeof(double));
    s->emissivity = (double *)xcalloc(s->N * s->L, sizeof(double));
    s->opacity  = (double *)xcalloc(s->N * s->L, sizeof(double));
    s->J        = (double *)xcalloc(s->N * s->F, sizeof(double));
    s->Ls       = (double *)xcalloc(s->N * s->L, sizeof(double));
    s->Jlin     = (double *)xcalloc(s->N * s->L, sizeof(double));
    s->Jeff     = (double *)xcalloc(s->N * s->L, sizeof(double));
    s->Jdif     = (double *)xcalloc(s->N * s->L, sizeof(double));
    if (m->n_off_diag >= 2 && s->S == 1) {
        /* at most two elements per band and ray pair (raypair_lambda.tpp:129-146) */
        s->band_cap = (s->R / 2) * 2 * (m->n_off_diag - 1);
        s->band_n   = (int64_t *)xcalloc(s->N * s->L, sizeof(int64_t));
        s->band_nr  = (int64_t *)xcalloc(s->N * s->L * s->band_cap, sizeof(int64_t));
        s->band_Ls  = (double *)xcalloc(s->N * s->L * s->band_cap, sizeof(double));
    }
    return s;
}

void orc_destroy(orc_state *s)
{
    if (!s) return;
    free(s->line); free(s->line_index); free(s->cell2boundary_nr); free(s->boundary);
    free(s->antipod); free(s->nb_off); free(s->nu); free(s->nr_line); free(s->appears);
    free(s->corr_l); free(s->corr_k); free(s->corr_z); free(s->I_bdy); free(s->pop);
    free(s->prev1); free(s->prev2); free(s->prev3); free(s->pop_tot); free(s->emissivity);
    free(s->opacity); free(s->J); free(s->Ls); free(s->Jlin); free(s->Jeff); free(s->Jdif);
    free(s->u); free(s->band_n); free(s->band_nr); free(s->band_Ls); free(s->sp);
    free(s);
}

/* ---- Functions ---------------------------------------------------------- */

/* planck (src/Functions/planck.tpp:15-24) */
double orc_planck(double temperature, double freq)
{
    const double h_over_kbT = HH_OVER_KB / temperature;
    return TWO_HH_OVER_CC_SQUARED * (freq*freq*freq) / expm1(h_over_kbT*freq);
}

/* Thermodynamics::profile(width, freq_diff) (thermodynamics.tpp:16-29) with the
 * scalar vExpMinus = 1/exp (src/Tools/Parallel/wrap_Grid.hpp:145-163). */
double orc_profile(double width, double freq_diff)
{
    const double inverse_width = 1.0 / width;
    const double sqrtExponent  = inverse_width * freq_diff;
    const double exponent      = sqrtExponent * sqrtExponent;
    return inverse_width * O_INVERSE_SQRT_PI * (1.0 / exp(exponent));
}

/* Thermodynamics::profile_width(inverse_mass, p) (thermodynamics.tpp:62-69) */
static double o_profile_width(const orc_state *s, double inverse_mass, int64_t p)
{
    return sqrt(TWO_KB_OVER_AMU_C_SQUARED * inverse_mass * s->m.temperature[p] + s->m.vturb2[p]);
}

/* ... of species 0 */
double orc_profile_width(const orc_state *s, int64_t p)
{
    return o_profile_width(s, s->sp[0].v.inverse_mass, p);
}

/* search (src/Functions/interpolation.tpp:14-54) */
static long o_search(const double *x, long n, double value)
{
    long start = 0, stop = n - 1;
    if (value >= x[stop]) return stop;
    else if (value <= x[start]) return start;
    while (stop > start + 1) {
        const long middle = (stop + start) / 2;
        if (value > x[middle]) start = middle;
        else if (value < x[middle]) stop = middle;
        else return middle;
    }
    return stop;
}

/* CollisionPartner::adjust_abundance_for_ortho_or_para and
 * ::interpolate_collision_coefficients (collisionPartner.tpp:15-72).
 * Partner c's tables start at the concatenated offsets. */
static void o_collision_rates(const mgr_species_view *m, int64_t c, double temperature,
                              double *Ce_intpld, double *Cd_intpld, double *abundance_factor)
{
    int64_t koff = 0, toff = 0, ctoff = 0;
    for (int64_t q = 0; q < c; q++) {
        koff  += m->colpar_ncol[q];
        toff  += m->colpar_ntmp[q];
        ctoff += m->colpar_ncol[q] * m->colpar_ntmp[q];
    }
    (void)koff;
    const int64_t ncol = m->colpar_ncol[c], ntmp = m->colpar_ntmp[c];
    const double *tmp = m->tmp + toff;
    const double *Ce = m->Ce + ctoff, *Cd = m->Cd + ctoff;

    double factor = 1.0;
    if (m->colpar_orth_or_para[c] != 'n') {
        const double frac_H2_para = 1.0 / (1.0 + 9.0*exp(-170.5/temperature));
        if (m->colpar_orth_or_para[c] == 'o') factor *= (1.0 - frac_H2_para);
        if (m->colpar_orth_or_para[c] == 'p') factor *= frac_H2_para;
    }
    *abundance_factor = factor;

    const long t = o_search(tmp, (long)ntmp, temperature);
    if (t == 0) {
        for (int64_t k = 0; k < ncol; k++) { Ce_intpld[k] = Ce[k]; Cd_intpld[k] = Cd[k]; }
    } else if (t == ntmp) {   /* unreachable: search returns at most ntmp-1 */
        for (int64_t k = 0; k < ncol; k++) {
            Ce_intpld[k] = Ce[(ntmp-1)*ncol + k]; Cd_intpld[k] = Cd[(ntmp-1)*ncol + k];
        }
    } else {
        const double step = (temperature - tmp[t-1]) / (tmp[t] - tmp[t-1]);
        for (int64_t k = 0; k < ncol; k++) {
            Ce_intpld[k] = Ce[(t-1)*ncol + k] + (Ce[t*ncol + k] - Ce[(t-1)*ncol + k]) * step;
            Cd_intpld[k] = Cd[(t-1)*ncol + k] + (Cd[t*ncol + k] - Cd[(t-1)*ncol + k]) * step;
        }
    }
}

/* ... for partner c of species 0 */
void orc_collision_rates(const orc_state *s, int64_t c, double temperature,
                         double *Ce_intpld, double *Cd_intpld, double *abundance_factor)
{
    o_collision_rates(&s->sp[0].v, c, temperature, Ce_intpld, Cd_intpld, abundance_factor);
}

/* ---- Model set-up ------------------------------------------------------- */

/* Simulation::compute_spectral_discretisation (src/Simulation/sim_spectrum.tpp:13-196) */
static int o_spectral_discretisation(orc_state *s, int image, double image_width);

int orc_compute_spectral_discretisation(orc_state *s) { return o_spectral_discretisation(s, 0, 0.0); }

/* Simulation::compute_spectral_discretisation_image (sim_spectrum.tpp:206-311):
 * the same construction with one fixed relative width for every cell. */
int orc_compute_spectral_discretisation_image(orc_state *s, double width)
{
    return o_spectral_discretisation(s, 1, width);
}

static int o_spectral_discretisation(orc_state *s, int image, double image_width)
{
    const int64_t F = s->F, L = s->L, Q = s->Q;
    #pragma omp parallel
    {
        vi_pair *fr = (vi_pair *)xcalloc(F, sizeof(vi_pair));
        int64_t *inv = (int64_t *)xcalloc(F, sizeof(int64_t));
        #pragma omp for schedule(static)
        for (int64_t p = 0; p < s->N; p++) {
            int64_t index0 = 0, index1 = 0;
            for (int64_t l = 0; l < s->S; l++) {
                const mgr_species_view *sv = &s->sp[l].v;
                const double pw = o_profile_width(s, sv->inverse_mass, p);
                for (int64_t k = 0; k < sv->nrad; k++) {
                    /* the sorted list at the running line index, with the width
                     * of species l (sim_spectrum.tpp:23-45) */
                    const double freqs_line = s->line[index0];
                    const double width = freqs_line * pw;
                    for (int64_t z = 0; z < Q; z++) {
                        if (image)   /* freqs_line + freqs_line * width * root (sim_spectrum.tpp:227) */
                            fr[index1].v = freqs_line + freqs_line * image_width * sv->roots[z];
                        else
                            fr[index1].v = freqs_line + width * sv->roots[z];
                        fr[index1].i = index1;
                        index1++;
                    }
                    index0++;
                }
            }
            qsort(fr, (size_t)F, sizeof(vi_pair), vi_cmp);
            for (int64_t fl = 0; fl < F; fl++) {
                s->nu[p*F + fl] = fr[fl].v;
                inv[fr[fl].i] = fl;
            }
            int64_t index2 = 0;
            for (int64_t k = 0; k < L; k++)
                for (int64_t z = 0; z < Q; z++) {
                    s->nr_line[(p*L + k)*Q + z] = inv[index2];
                    index2++;
                }
        }
        free(fr); free(inv);
    }
    /* cell-independent back-maps, indexed by the unsorted position (P11) */
    for (int64_t fl = 0; fl < F; fl++) {
        s->appears[fl] = 0; s->corr_l[fl] = F; s->corr_k[fl] = F; s->corr_z[fl] = F;
    }
    int64_t index2 = 0;
    for (int64_t l = 0; l < s->S; l++)
        for (int64_t k = 0; k < s->sp[l].L; k++)
            for (int64_t z = 0; z < Q; z++) {
                s->appears[index2] = 1; s->corr_l[index2] = l;
                s->corr_k[index2] = k;  s->corr_z[index2] = z;
                index2++;
            }
    s->spec_set = 1;
    return 0;
}

/* Simulation::compute_boundary_intensities (src/Simulation/sim_radiation.tpp:17-48) */
int orc_compute_boundary_intensities(orc_state *s, const double *temperatures)
{
    const int64_t F = s->F;
    for (int64_t r = 0; r < s->R / 2; r++)
        for (int64_t b = 0; b < s->NB; b++) {
            const int64_t p = s->m.boundary2cell_nr[b];
            const double T = temperatures ? temperatures[b] : O_T_CMB;
            for (int64_t f = 0; f < F; f++)
                s->I_bdy[(r*s->NB + b)*F + f] = orc_planck(T, s->nu[p*F + f]);
        }
    return 0;
}

/* LineProducingSpecies::get_emissivity / get_opacity (lineProducingSpecies.tpp:41-70)
 * and Lines::set_emissivity_and_opacity (lines.tpp:44-60). */
void orc_set_emissivity_and_opacity(orc_state *s)
{
    const int64_t L = s->L;
    for (int64_t l = 0; l < s->S; l++) {
        const mgr_species_view *sv = &s->sp[l].v;
        const int64_t V = s->sp[l].V, loff = s->sp[l].loff;
        const double *pop = s->pop + s->N * s->sp[l].voff;
        for (int64_t p = 0; p < s->N; p++)
            for (int64_t k = 0; k < s->sp[l].L; k++) {
                const int64_t i = sv->irad[k], j = sv->jrad[k];
                s->emissivity[p*L + loff + k] = HH_OVER_FOUR_PI * sv->A[k] * pop[p*V + i];
                s->opacity[p*L + loff + k] = HH_OVER_FOUR_PI * (pop[p*V + j] * sv->Ba[k]
                                                                - pop[p*V + i] * sv->Bs[k]);
            }
    }
}

/* LineProducingSpecies::update_using_LTE (lineProducingSpecies.tpp:82-115) */
int orc_compute_LTE_level_populations(orc_state *s)
{
    for (int64_t l = 0; l < s->S; l++) {
        const mgr_species_view *sv = &s->sp[l].v;
        const int64_t V = s->sp[l].V;
        double *pop = s->pop + s->N * s->sp[l].voff, *pop_tot = s->pop_tot + l * s->N;
        for (int64_t p = 0; p < s->N; p++) {
            pop_tot[p] = s->m.abundance[p*s->m.nspecs + sv->num];
            double pf = 0.0;
            for (int64_t i = 0; i < V; i++) {
                pop[p*V + i] = sv->weight[i] * exp(-sv->energy[i] / (O_KB * s->m.temperature[p]));
                pf += pop[p*V + i];
            }
            for (int64_t i = 0; i < V; i++) pop[p*V + i] *= pop_tot[p] / pf;
        }
    }
    orc_set_emissivity_and_opacity(s);
    return 0;
}

/* ---- Geometry ----------------------------------------------------------- */

/* Simulation::get_dshift_max (src/Simulation/simulation.tpp:15-37): the
 * smallest over the species */
double orc_get_dshift_max(const orc_state *s, int64_t o)
{
    double dshift_max = DBL_MAX;
    for (int64_t l = 0; l < s->S; l++) {
        const double nd = s->m.max_width_fraction * o_profile_width(s, s->sp[l].v.inverse_mass, o);
        if (dshift_max > nd) dshift_max = nd;
    }
    return dshift_max;
}

/* Geometry::get_doppler_shift (geometry.tpp:433-465); Eigen's Vector3d dot is
 * evaluated as (x*x' + y*y') + z*z' (SSE2 packet of two, then the tail). */
double orc_get_doppler_shift(const orc_state *s, int64_t o, int64_t r, int64_t c, int rest_frame)
{
    const double *v = s->m.velocity;
    if (!rest_frame) {
        const double *n = s->m.rays + 3*r;
        const double dx = v[3*c+0] - v[3*o+0];
        const double dy = v[3*c+1] - v[3*o+1];
        const double dz = v[3*c+2] - v[3*o+2];
        return 1.0 - ((dx*n[0] + dy*n[1]) + dz*n[2]);
    } else {
        int64_t rc = r;
        if (r >= s->R / 2) rc = s->antipod[r];
        const double *n = s->m.rays + 3*rc;
        return 1.0 - ((v[3*c+0]*n[0] + v[3*c+1]*n[1]) + v[3*c+2]*n[2]);
    }
}

/* Geometry::get_next_general (geometry.tpp:328-421) */
static int64_t o_get_next(const orc_state *s, int64_t origin, int64_t ray, int64_t current,
                          double *Z, double *dZ)
{
    const double *pos = s->m.position;
    const double *n = s->m.rays + 3*ray;
    double dmin = DBL_MAX;
    int64_t next = -1;
    for (int64_t q = s->nb_off[current]; q < s->nb_off[current+1]; q++) {
        const int64_t nb = s->m.neighbors[q];
        const double Rx = pos[3*nb+0] - pos[3*origin+0];
        const double Ry = pos[3*nb+1] - pos[3*origin+1];
        const double Rz = pos[3*nb+2] - pos[3*origin+2];
        const double Z_new = (Rx*n[0] + Ry*n[1]) + Rz*n[2];
        if (Z_new > *Z) {
            const double distance_from_ray2 = ((Rx*Rx + Ry*Ry) + Rz*Rz) - Z_new*Z_new;
            if (distance_from_ray2 < dmin) {
                dmin = distance_from_ray2;
                next = nb;
                *dZ = Z_new - *Z;
            }
        }
    }
    *Z = *Z + *dZ;
    return next;
}

/* Geometry::set_data (geometry.tpp:100-207).  Appends to the output arrays;
 * returns the new count (unchanged when more than 10000 points are needed).
 * o_crt receives crt for a plain point and -2 - step (the index of the traced
 * step) for interpolation points, which the reference leaves unset. */
static int64_t o_set_data(int64_t crt, int64_t nxt, double shift_crt, double shift_nxt,
                          double dZ_loc, double dshift_max, int64_t cnt, int64_t cap,
                          int64_t *o_crt, int64_t *o_cell, double *o_shift, double *o_dZ,
                          int64_t step)
{
    const double dshift     = shift_nxt - shift_crt;
    const double dshift_abs = fabs(dshift);
    if (dshift_abs > dshift_max) {
        const long   n_interpl      = (long)(dshift_abs / dshift_max + 1);
        const long   half_n_interpl = (long)(0.5 * n_interpl);
        const double dZ_interpl     = dZ_loc / n_interpl;
        const double dshift_interpl = dshift / n_interpl;
        if ((n_interpl > 10000) || (n_interpl < 0)) return cnt;
        for (long m = 1; m < half_n_interpl; m++) {
            if (cnt < cap) {
                if (o_crt) o_crt[cnt] = -2 - step;
                o_cell[cnt] = crt; o_shift[cnt] = shift_crt + m * dshift_interpl; o_dZ[cnt] = dZ_interpl;
            }
            cnt++;
        }
        for (long m = half_n_interpl; m <= n_interpl; m++) {
            if (cnt < cap) {
                if (o_crt) o_crt[cnt] = -2 - step;
                o_cell[cnt] = nxt; o_shift[cnt] = shift_crt + m * dshift_interpl; o_dZ[cnt] = dZ_interpl;
            }
            cnt++;
        }
    } else {
        if (cnt < cap) {
            if (o_crt) o_crt[cnt] = crt;
            o_cell[cnt] = nxt; o_shift[cnt] = shift_nxt; o_dZ[cnt] = dZ_loc;
        }
        cnt++;
    }
    return cnt;
}

static int64_t o_trace_ray(const orc_state *s, int64_t o, int64_t r, double dshift_max, int64_t cap,
                           int64_t *crt_out, int64_t *cellNr, double *shift, double *dZs, int rest);

/* Geometry::trace_ray<CoMoving> (geometry.tpp:21-95) */
int64_t orc_trace_ray(const orc_state *s, int64_t o, int64_t r, double dshift_max,
                      int64_t cap, int64_t *crt_out, int64_t *cellNr, double *shift, double *dZs)
{
    return o_trace_ray(s, o, r, dshift_max, cap, crt_out, cellNr, shift, dZs, 0);
}

/* Geometry::trace_ray<Rest> (geometry.tpp:21-95 with the rest-frame shift) */
int64_t orc_trace_ray_rest(const orc_state *s, int64_t o, int64_t r, double dshift_max,
                           int64_t cap, int64_t *crt_out, int64_t *cellNr, double *shift, double *dZs)
{
    return o_trace_ray(s, o, r, dshift_max, cap, crt_out, cellNr, shift, dZs, 1);
}

static int64_t o_trace_ray(const orc_state *s, int64_t o, int64_t r, double dshift_max, int64_t cap,
                           int64_t *crt_out, int64_t *cellNr, double *shift, double *dZs, int rest)
{
    double Z = 0.0, dZ = 0.0;
    int64_t cnt = 0;
    int64_t nxt = o_get_next(s, o, r, o, &Z, &dZ);
    if (nxt != -1) {
        int64_t crt = o;
        double shift_crt = orc_get_doppler_shift(s, o, r, crt, rest);
        double shift_nxt = orc_get_doppler_shift(s, o, r, nxt, rest);
        int64_t step = 0;
        cnt = o_set_data(crt, nxt, shift_crt, shift_nxt, dZ, dshift_max, cnt, cap,
                         crt_out, cellNr, shift, dZs, step);
        while (!s->boundary[nxt]) {
            shift_crt = shift_nxt;
            crt = nxt;
            nxt = o_get_next(s, o, r, nxt, &Z, &dZ);
            if (nxt < 0) return -1;
            shift_nxt = orc_get_doppler_shift(s, o, r, nxt, rest);
            step++;
            cnt = o_set_data(crt, nxt, shift_crt, shift_nxt, dZ, dshift_max, cnt, cap,
                             crt_out, cellNr, shift, dZs, step);
        }
    }
    return cnt;
}

/* get_required_npoints / Geometry::get_npoints_on_ray (geometry.tpp:470-543) */
int64_t orc_get_npoints_on_ray(const orc_state *s, int64_t o, int64_t r, double dshift_max)
{
    int64_t npoints = 0;
    double Z = 0.0, dZ = 0.0;
    int64_t nxt = o_get_next(s, o, r, o, &Z, &dZ);
    if (nxt != -1) {
        double shift_crt = orc_get_doppler_shift(s, o, r, o, 0);
        double shift_nxt = orc_get_doppler_shift(s, o, r, nxt, 0);
        for (;;) {
            const double dshift_abs = fabs(shift_nxt - shift_crt);
            npoints += (dshift_abs > dshift_max) ? (int64_t)(dshift_abs/dshift_max + 1) : 1;
            if (s->boundary[nxt]) break;
            shift_crt = shift_nxt;
            nxt = o_get_next(s, o, r, nxt, &Z, &dZ);
            if (nxt < 0) return -1;
            shift_nxt = orc_get_doppler_shift(s, o, r, nxt, 0);
        }
    }
    return npoints;
}

/* ---- Line emissivity and opacity ---------------------------------------- */

/* Simulation::get_line_width (src/Simulation/simulation.tpp:363-373): the
 * width of sorted line lindex, with the inverse mass of the species that the
 * frequency back-map names at index line_index[lindex] (a bin index there). */
static double o_line_width(const orc_state *s, int64_t p, int64_t lindex)
{
    const int64_t f = s->line_index[lindex];
    const int64_t l = s->corr_l[f] < s->S ? s->corr_l[f] : 0;
    return s->line[lindex] * o_profile_width(s, s->sp[l].v.inverse_mass, p);
}

/* Simulation::get_eta_and_chi (src/Simulation/simulation.tpp:385-481).
 * lnotch is stateful. */
void orc_get_eta_and_chi(const orc_state *s, double freq_scaled, int64_t p,
                         int64_t *lnotch, double *eta, double *chi)
{
    const int64_t L = s->L;
    *eta = 0.0;
    *chi = 1.0E-26;
    const double lower = -11.0, upper = +11.0;
    double freq_diff = freq_scaled - s->line[*lnotch];
    double width = o_line_width(s, p, *lnotch);
    while (freq_diff > upper*width) {
        if (*lnotch >= L-1) break;
        (*lnotch)++;
        freq_diff = freq_scaled - s->line[*lnotch];
        width = o_line_width(s, p, *lnotch);
    }
    int64_t lindex = *lnotch;
    while (freq_diff >= lower*width) {
        const double line_profile = orc_profile(width, freq_diff);
        const int64_t ind = s->line_index[lindex] + p*L;
        *eta += s->emissivity[ind] * freq_scaled * line_profile;
        *chi += s->opacity[ind] * freq_scaled * line_profile;
        if (lindex >= L-1) break;
        lindex++;
        freq_diff = freq_scaled - s->line[lindex];
        width = o_line_width(s, p, lindex);
    }
}

/* Radiation::rescale_I_bdy, scalar branch (src/Model/Radiation/radiation.tpp:341-373),
 * with search_with_notch from notch 0 (interpolation.tpp:96-110) and
 * interpolate_linear (interpolation.tpp:126-134). */
static double o_rescale_I_bdy(const orc_state *s, double freq_scaled, int64_t R, int64_t p,
                              int64_t b)
{
    const int64_t F = s->F;
    const double *nup = s->nu + p*F;
    const double *Ib = s->I_bdy + (R*s->NB + b)*F;
    int64_t notch = 0;
    while (notch < F) { if (freq_scaled <= nup[notch]) break; notch++; }
    if (notch >= F) notch = F - 1;
    if ((notch == 0) || (notch == F-1)) return Ib[notch];
    const double nu1 = nup[notch], nu2 = nup[notch-1];
    const double I1 = Ib[notch], I2 = Ib[notch-1];
    return (I2 - I1) / (nu2 - nu1) * (freq_scaled - nu1) + I1;
}

double orc_rescale_I_bdy(const orc_state *s, double freq_scaled, int64_t R, int64_t p, int64_t b)
{
    return o_rescale_I_bdy(s, freq_scaled, R, p, b);
}

int64_t orc_cell2boundary_nr(const orc_state *s, int64_t p) { return s->cell2boundary_nr[p]; }

/* ---- Feautrier solve ---------------------------------------------------- */

/* RayPair::solve (src/Simulation/Raypair/raypair_solver.tpp:200-420) */
void orc_raypair_solve(int64_t ndep, int64_t n_ar, int64_t first, int64_t last,
                       int64_t n_off_diag, const double *term1, const double *dtau,
                       double I_bdy_0, double I_bdy_n,
                       double *Su, double *L_diag, double *L_upper, double *L_lower,
                       double *A_out, double *C_out)
{
    double *A  = (double *)xcalloc(ndep, sizeof(double));
    double *C  = (double *)xcalloc(ndep, sizeof(double));
    double *F  = (double *)xcalloc(ndep, sizeof(double));
    double *G  = (double *)xcalloc(ndep, sizeof(double));
    double *iA = (double *)xcalloc(ndep, sizeof(double));
    double *iC = (double *)xcalloc(ndep, sizeof(double));
    double *i1F = (double *)xcalloc(ndep, sizeof(double));
    double *i1G = (double *)xcalloc(ndep, sizeof(double));
    double *GoG = (double *)xcalloc(ndep, sizeof(double));

    const double inverse_dtau0 = 1.0 / dtau[first];
    const double inverse_dtaud = 1.0 / dtau[last-1];
    C[first] = 2.0 * inverse_dtau0 * inverse_dtau0;
    const double B0_min_C0 = 1.0 + 2.0 * inverse_dtau0;
    const double B0 = B0_min_C0 + C[first];
    const double inverse_B0 = 1.0 / B0;
    Su[first] = term1[first] + 2.0 * I_bdy_0 * inverse_dtau0;
    for (int64_t n = first+1; n < last; n++) {
        iA[n] = 0.5 * (dtau[n-1] + dtau[n]) * dtau[n-1];
        A[n]  = 1.0 / iA[n];
        iC[n] = 0.5 * (dtau[n-1] + dtau[n]) * dtau[n];
        C[n]  = 1.0 / iC[n];
        Su[n] = term1[n];
    }
    A[last] = 2.0 * inverse_dtaud * inverse_dtaud;
    const double Bd_min_Ad = 1.0 + 2.0 * inverse_dtaud;
    const double Bd = Bd_min_Ad + A[last];
    Su[last] = term1[last] + 2.0 * I_bdy_n * inverse_dtaud;

    /* elimination */
    Su[first] = Su[first] * inverse_B0;
    F[first] = 0.5 * B0_min_C0 * dtau[first] * dtau[first];
    i1F[first] = 1.0 / (1.0 + F[first]);
    for (int64_t n = first+1; n < last; n++) {
        F[n] = (1.0 + A[n]*F[n-1]*i1F[n-1]) * iC[n];
        i1F[n] = 1.0 / (1.0 + F[n]);
        Su[n] = (Su[n] + A[n]*Su[n-1]) * i1F[n] * iC[n];
    }
    const double denominator = 1.0 / (Bd_min_Ad + Bd*F[last-1]);
    Su[last] = (Su[last] + A[last]*Su[last-1]) * (1.0 + F[last-1]) * denominator;

    if (n_off_diag == 0) {
        if (n_ar < last) {
            G[last] = 0.5 * Bd_min_Ad * dtau[last-1] * dtau[last-1];
            GoG[last] = G[last] / (1.0 + G[last]);
            for (int64_t n = last-1; n > n_ar; n--) {
                Su[n] = Su[n] + Su[n+1] * i1F[n];
                G[n] = (1.0 + C[n]*GoG[n+1]) * iA[n];
                GoG[n] = G[n] / (1.0 + G[n]);
            }
            Su[n_ar] = Su[n_ar] + Su[n_ar+1] * i1F[n_ar];
            L_diag[n_ar] = iC[n_ar] / (F[n_ar] + GoG[n_ar+1]);
        } else {
            L_diag[last] = (1.0 + F[last-1]) / (Bd_min_Ad + Bd*F[last-1]);
        }
    } else {
        G[last] = 0.5 * Bd_min_Ad * dtau[last-1] * dtau[last-1];
        i1G[last] = 1.0 / (1.0 + G[last]);
        for (int64_t n = last-1; n > first; n--) {
            Su[n] = Su[n] + Su[n+1] * i1F[n];
            G[n] = (1.0 + C[n]*G[n+1]*i1G[n+1]) * iA[n];
            i1G[n] = 1.0 / (1.0 + G[n]);
        }
        Su[first] = Su[first] + Su[first+1] * i1F[first];
        L_diag[last] = (1.0 + F[last-1]) / (Bd_min_Ad + Bd*F[last-1]);
        for (int64_t n = last-1; n >= first+1; n--)
            L_diag[n] = iC[n] / (F[n] + G[n+1] * i1G[n+1]);
        L_diag[first] = (1.0 + G[first+1]) / (B0_min_C0 + B0*G[first+1]);
        for (int64_t n = last-1; n >= first; n--) {
            L_upper[0*ndep + n] = L_diag[n+1] * i1F[n];
            L_lower[0*ndep + n] = L_diag[n] * i1G[n+1];
        }
        for (int64_t m = 1; m < n_off_diag; m++)
            for (int64_t n = last-1-m; n >= first; n--) {
                L_upper[m*ndep + n] = L_upper[(m-1)*ndep + n+1] * i1F[n];
                L_lower[m*ndep + n] = L_lower[(m-1)*ndep + n] * i1G[n+m+1];
            }
    }
    if (A_out) for (int64_t n = 0; n < ndep; n++) A_out[n] = A[n];
    if (C_out) for (int64_t n = 0; n < ndep; n++) C_out[n] = C[n];
    free(A); free(C); free(F); free(G); free(iA); free(iC); free(i1F); free(i1G); free(GoG);
}

/* ---- Radiation field ---------------------------------------------------- */

typedef struct {
    int64_t cap, nband;   /* nband: rows of L_upper / L_lower */
    int64_t *cell_r, *cell_ar; double *shift_r, *shift_ar, *dZ_r, *dZ_ar;
    int64_t *lnotch_r, *lnotch_ar;
    double *term1, *chi, *frs, *dtau, *Su, *L_diag, *L_upper, *L_lower;
    int64_t *nrs;
} o_pair_buf;

static void o_buf_alloc(o_pair_buf *b, int64_t cap, int64_t nband)
{
    b->cap = cap;
    b->nband = nband > 1 ? nband : 1;
    b->cell_r = (int64_t *)xcalloc(cap, 8); b->cell_ar = (int64_t *)xcalloc(cap, 8);
    b->shift_r = (double *)xcalloc(cap, 8); b->shift_ar = (double *)xcalloc(cap, 8);
    b->dZ_r = (double *)xcalloc(cap, 8);    b->dZ_ar = (double *)xcalloc(cap, 8);
    b->lnotch_r = (int64_t *)xcalloc(cap, 8); b->lnotch_ar = (int64_t *)xcalloc(cap, 8);
    const int64_t d = 2*cap + 1;
    b->term1 = (double *)xcalloc(d, 8); b->chi = (double *)xcalloc(d, 8);
    b->frs = (double *)xcalloc(d, 8);   b->dtau = (double *)xcalloc(d, 8);
    b->Su = (double *)xcalloc(d, 8);    b->L_diag = (double *)xcalloc(d, 8);
    b->L_upper = (double *)xcalloc(b->nband * d, 8); b->L_lower = (double *)xcalloc(b->nband * d, 8);
    b->nrs = (int64_t *)xcalloc(d, 8);
}
static void o_buf_free(o_pair_buf *b)
{
    free(b->cell_r); free(b->cell_ar); free(b->shift_r); free(b->shift_ar); free(b->dZ_r);
    free(b->dZ_ar); free(b->lnotch_r); free(b->lnotch_ar); free(b->term1); free(b->chi);
    free(b->frs); free(b->dtau); free(b->Su); free(b->L_diag); free(b->nrs);
    free(b->L_upper); free(b->L_lower);
}

/* Simulation::setup (src/Simulation/simulation.tpp:185-353): term1, chi and
 * dtau of the pair in b, the solved range [first, last] and the boundary
 * intensities at both ends. */
static void o_setup(const orc_state *s, int64_t R, int64_t origin, int64_t f,
                    int64_t n_ar, const int64_t *c_ar, const double *s_ar, const double *z_ar,
                    int64_t *ln_ar, int64_t n_r, const int64_t *c_r, const double *s_r,
                    const double *z_r, int64_t *ln_r, int64_t *lnotch_at_origin, o_pair_buf *b,
                    int64_t *first_out, int64_t *last_out, double *I0_out, double *In_out)
{
    const int64_t F = s->F;
    int64_t first = 0, last = n_ar;
    double eta, chi, I_bdy_0, I_bdy_n;
    double freq_scaled = s->nu[origin*F + f];
    int64_t index = n_ar;

    orc_get_eta_and_chi(s, freq_scaled, origin, lnotch_at_origin, &eta, &chi);
    b->term1[index] = eta / chi;
    b->chi[index] = chi; b->nrs[index] = origin; b->frs[index] = freq_scaled;
    double chi_prev = chi;
    const double chi_o = chi;

    if (n_ar > 0) {
        index = n_ar - 1;
        for (int64_t i = 0; i < n_ar; i++) {
            freq_scaled = s_ar[i] * s->nu[origin*F + f];
            orc_get_eta_and_chi(s, freq_scaled, c_ar[i], &ln_ar[i], &eta, &chi);
            b->term1[index] = eta / chi;
            b->dtau[index] = 0.5 * (chi + chi_prev) * z_ar[i];
            b->chi[index] = chi; b->nrs[index] = c_ar[i]; b->frs[index] = freq_scaled;
            chi_prev = chi;
            index--;
        }
        const int64_t cellNr = c_ar[n_ar-1];
        I_bdy_0 = o_rescale_I_bdy(s, freq_scaled, R, cellNr, s->cell2boundary_nr[cellNr]);
        first = index + 1;
        chi_prev = chi_o;
    } else {
        I_bdy_0 = s->I_bdy[(R*s->NB + s->cell2boundary_nr[origin])*F + f];
    }

    if (n_r > 0) {
        index = n_ar + 1;
        for (int64_t i = 0; i < n_r; i++) {
            freq_scaled = s_r[i] * s->nu[origin*F + f];
            orc_get_eta_and_chi(s, freq_scaled, c_r[i], &ln_r[i], &eta, &chi);
            b->term1[index] = eta / chi;
            b->dtau[index-1] = 0.5 * (chi + chi_prev) * z_r[i];
            b->chi[index] = chi; b->nrs[index] = c_r[i]; b->frs[index] = freq_scaled;
            chi_prev = chi;
            index++;
        }
        const int64_t cellNr = c_r[n_r-1];
        I_bdy_n = o_rescale_I_bdy(s, freq_scaled, R, cellNr, s->cell2boundary_nr[cellNr]);
        last = index - 1;
    } else {
        I_bdy_n = s->I_bdy[(R*s->NB + s->cell2boundary_nr[origin])*F + f];
    }

    *first_out = first;
    *last_out = last;
    *I0_out = I_bdy_0;
    *In_out = I_bdy_n;
}

/* Simulation::setup followed by RayPair::solve with the model's n_off_diag
 * (0; >= 1: the back-substitution branch, raypair_solver.tpp:336-388, which
 * also fills the bands).  Side "ar" is the longer ray.  Returns u at the
 * origin. */
static double o_setup_and_solve(const orc_state *s, int64_t R, int64_t origin, int64_t f,
                                int64_t n_ar, const int64_t *c_ar, const double *s_ar,
                                const double *z_ar, int64_t *ln_ar,
                                int64_t n_r, const int64_t *c_r, const double *s_r,
                                const double *z_r, int64_t *ln_r,
                                int64_t *lnotch_at_origin, o_pair_buf *b, double *L_origin,
                                double *chi_origin)
{
    const int64_t ndep = n_ar + n_r + 1;
    int64_t first, last;
    double I_bdy_0, I_bdy_n;
    o_setup(s, R, origin, f, n_ar, c_ar, s_ar, z_ar, ln_ar, n_r, c_r, s_r, z_r, ln_r, lnotch_at_origin, b,
            &first, &last, &I_bdy_0, &I_bdy_n);
    orc_raypair_solve(ndep, n_ar, first, last, s->m.n_off_diag, b->term1, b->dtau, I_bdy_0, I_bdy_n,
                      b->Su, b->L_diag, b->L_upper, b->L_lower, NULL, NULL);
    *L_origin = b->L_diag[n_ar];
    *chi_origin = b->chi[n_ar];
    return b->Su[n_ar];
}

/* RayPair::solve_for_image (src/Simulation/Raypair/raypair_solver.tpp:523-579):
 * the same elimination with back-substitution over the whole pair, so that
 * Su holds u at every point from 0 to last. */
void orc_solve_for_image(int64_t ndep, int64_t first, int64_t last, const double *term1,
                         const double *dtau, double I_bdy_0, double I_bdy_n, double *Su)
{
    double *A = (double *)xcalloc(ndep, sizeof(double));
    double *C = (double *)xcalloc(ndep, sizeof(double));
    double *F = (double *)xcalloc(ndep, sizeof(double));
    double *iA = (double *)xcalloc(ndep, sizeof(double));
    double *iC = (double *)xcalloc(ndep, sizeof(double));
    double *i1F = (double *)xcalloc(ndep, sizeof(double));
    const double inverse_dtau0 = 1.0 / dtau[first];
    const double inverse_dtaud = 1.0 / dtau[last-1];
    C[first] = 2.0 * inverse_dtau0 * inverse_dtau0;
    const double B0_min_C0 = 1.0 + 2.0 * inverse_dtau0;
    const double B0 = B0_min_C0 + C[first];
    const double inverse_B0 = 1.0 / B0;
    Su[first] = term1[first] + 2.0 * I_bdy_0 * inverse_dtau0;
    for (int64_t n = first+1; n < last; n++) {
        iA[n] = 0.5 * (dtau[n-1] + dtau[n]) * dtau[n-1];
        iC[n] = 0.5 * (dtau[n-1] + dtau[n]) * dtau[n];
        A[n] = 1.0 / iA[n];
        C[n] = 1.0 / iC[n];
        Su[n] = term1[n];
    }
    A[last] = 2.0 * inverse_dtaud * inverse_dtaud;
    const double Bd_min_Ad = 1.0 + 2.0 * inverse_dtaud;
    const double Bd = Bd_min_Ad + A[last];
    Su[last] = term1[last] + 2.0 * I_bdy_n * inverse_dtaud;
    Su[first] = Su[first] * inverse_B0;
    F[first] = 0.5 * B0_min_C0 * dtau[first] * dtau[first];
    i1F[first] = 1.0 / (1.0 + F[first]);
    for (int64_t n = first+1; n < last; n++) {
        F[n] = (1.0 + A[n]*F[n-1]*i1F[n-1]) * iC[n];
        i1F[n] = 1.0 / (1.0 + F[n]);
        Su[n] = (Su[n] + A[n]*Su[n-1]) * i1F[n] * iC[n];
    }
    const double denominator = 1.0 / (Bd_min_Ad + Bd*F[last-1]);
    Su[last] = (Su[last] + A[last]*Su[last-1]) * (1.0 + F[last-1]) * denominator;
    for (int64_t n = last-1; n >= 0; n--) Su[n] = Su[n] + Su[n+1] * i1F[n];
    free(A); free(C); free(F); free(iA); free(iC); free(i1F);
}

/* Simulation::compute_and_write_image (src/Simulation/sim_radiation.tpp:345-478)
 * for ray r < nrays/2: rest-frame traces of r and its antipode (no longer-ray
 * swap), u over the whole pair, and the intensities leaving it,
 * I_m[o][f] = 2 u_0 - I_bdy_0 (RayPair::get_Im_at_front) and
 * I_p[o][f] = 2 u_last - I_bdy_n (RayPair::get_Ip_at_end); a pair of one point
 * takes I_bdy of the origin's boundary entry.  Returns -1 on a trace error. */
int orc_compute_image(orc_state *s, int64_t r, double *I_m, double *I_p)
{
    const int64_t F = s->F;
    if (r < 0 || r >= s->R / 2) return -1;
    const int64_t R = r;
    int err = 0;
    #pragma omp parallel reduction(|:err)
    {
        o_pair_buf b;
        int64_t cap = 256;
        o_buf_alloc(&b, cap, 1);
        #pragma omp for schedule(static)
        for (int64_t o = 0; o < s->N; o++) {
            const int64_t ar = s->antipod[r];
            const double dshift_max = orc_get_dshift_max(s, o);
            int64_t n_r, n_ar;
            for (;;) {
                n_r = orc_trace_ray_rest(s, o, r, dshift_max, b.cap, NULL, b.cell_r, b.shift_r, b.dZ_r);
                n_ar = orc_trace_ray_rest(s, o, ar, dshift_max, b.cap, NULL, b.cell_ar, b.shift_ar, b.dZ_ar);
                if (n_r > b.cap || n_ar > b.cap) {
                    cap = 2 * (n_r > n_ar ? n_r : n_ar);
                    o_buf_free(&b); o_buf_alloc(&b, cap, 1);
                    continue;
                }
                break;
            }
            if (n_r < 0 || n_ar < 0) { err = 1; continue; }
            for (int64_t i = 0; i < n_r; i++) b.lnotch_r[i] = 0;
            for (int64_t i = 0; i < n_ar; i++) b.lnotch_ar[i] = 0;
            const int64_t ndep = n_ar + n_r + 1;
            if (ndep > 1) {
                int64_t lnotch_at_origin = 0;
                for (int64_t f = 0; f < F; f++) {
                    int64_t first, last;
                    double I0, In;
                    o_setup(s, R, o, f, n_ar, b.cell_ar, b.shift_ar, b.dZ_ar, b.lnotch_ar, n_r, b.cell_r,
                            b.shift_r, b.dZ_r, b.lnotch_r, &lnotch_at_origin, &b, &first, &last, &I0, &In);
                    orc_solve_for_image(ndep, first, last, b.term1, b.dtau, I0, In, b.Su);
                    I_m[o*F + f] = 2.0 * b.Su[0] - I0;
                    I_p[o*F + f] = 2.0 * b.Su[ndep-1] - In;
                }
            } else {
                const int64_t bb = s->cell2boundary_nr[o];
                for (int64_t f = 0; f < F; f++) {
                    I_m[o*F + f] = s->I_bdy[(R*s->NB + bb)*F + f];
                    I_p[o*F + f] = s->I_bdy[(R*s->NB + bb)*F + f];
                }
            }
        }
        o_buf_free(&b);
    }
    return err ? -1 : 0;
}

/* Simulation::compute_radiation_field (src/Simulation/sim_radiation.tpp:56-338),
 * RayPair branch, n_off_diag = 0, with RayPair::update_Lambda
 * (raypair_lambda.tpp:94-149) and Lambda::add_element (lambda.tpp:130-143). */
static int o_radiation_field(orc_state *s, int nthreads, const int64_t *list, int64_t o_begin,
                             int64_t o_end);

/* Radiation::u (radiation.cpp:117-126, sim_radiation.tpp:226-237, 258-262):
 * with scattering every compute_radiation_field keeps each ray pair's u at
 * the origin, u[R][f + p*F]. */
void orc_set_scattering(orc_state *s, int use_scattering)
{
    free(s->u);
    s->u = use_scattering ? (double *)xcalloc((size_t)(s->R / 2) * s->N * s->F, sizeof(double)) : NULL;
}

void orc_get_u(const orc_state *s, double *u)
{
    if (s->u) memcpy(u, s->u, (size_t)(s->R / 2) * s->N * s->F * sizeof(double));
}

int orc_compute_radiation_field(orc_state *s, int nthreads, int64_t o_begin, int64_t o_end)
{
    if (s->m.n_off_diag >= 2 && s->S > 1) return -1;   /* bands are restated for one species */
    if (o_begin < 0) o_begin = 0;
    if (o_end > s->N || o_end < 0) o_end = s->N;
    return o_radiation_field(s, nthreads, NULL, o_begin, o_end);
}

int orc_compute_radiation_field_list(orc_state *s, int nthreads, const int64_t *origins, int64_t n)
{
    if (s->m.n_off_diag >= 2 && s->S > 1) return -1;
    return o_radiation_field(s, nthreads, origins, 0, n);
}

/* Lambda::add_element (lambda.tpp:130-143): an element of (p, k) from cell
 * nr adds to the element of the same cell if there is one, else it is
 * appended.  The element nr = p is held in Ls. */
static void o_add_element(orc_state *s, int64_t p, int64_t k, int64_t nr, double Lel)
{
    if (nr == p) { s->Ls[p*s->L + k] += Lel; return; }
    const int64_t pk = p*s->L + k;
    int64_t *nrs = s->band_nr + pk*s->band_cap;
    double *Ls = s->band_Ls + pk*s->band_cap;
    for (int64_t e = 0; e < s->band_n[pk]; e++)
        if (nrs[e] == nr) { Ls[e] += Lel; return; }
    nrs[s->band_n[pk]] = nr;
    Ls[s->band_n[pk]] = Lel;
    s->band_n[pk]++;
}

/* RayPair::update_Lambda bands m = 1 .. n_off_diag-1 (raypair_lambda.tpp:129-146,
 * with get_L_lower / get_L_upper, raypair_lambda.tpp:44-84) for origin o and
 * bin f of transition k; the pair was solved with the longer ray (nA points)
 * on the low side.  Only elements this solve wrote are added (see the header). */
static void o_update_bands(orc_state *s, int64_t o, int64_t k, double factor, double freq_line,
                           int64_t nA, int64_t nB, const o_pair_buf *b)
{
    const int64_t ndep = nA + nB + 1;
    for (int64_t m = 1; m < s->m.n_off_diag; m++) {
        if (nA - m >= 0 && nB >= 1) {
            const int64_t n = nA - m;
            const double prof = orc_profile(freq_line * orc_profile_width(s, b->nrs[n]), b->frs[n] - freq_line);
            const double Lel = factor * (b->frs[n] * prof * b->L_lower[m*ndep + n] / b->chi[n]);
            o_add_element(s, o, k, b->nrs[n], Lel);
        }
        if (nA + m < ndep - m && nB > 2*m) {
            const int64_t n = nA + m;
            const double prof = orc_profile(freq_line * orc_profile_width(s, b->nrs[n]), b->frs[n] - freq_line);
            const double Lel = factor * (b->frs[n] * prof * b->L_upper[m*ndep + n] / b->chi[n]);
            o_add_element(s, o, k, b->nrs[n], Lel);
        }
    }
}

/* list == NULL: origins o_begin..o_end-1; otherwise origins list[o_begin..o_end-1] */
static int o_radiation_field(orc_state *s, int nthreads, const int64_t *list, int64_t o_begin,
                             int64_t o_end)
{
    const int64_t F = s->F, L = s->L;
    for (int64_t i = o_begin; i < o_end; i++) {
        const int64_t p = list ? list[i] : i;
        for (int64_t f = 0; f < F; f++) s->J[p*F + f] = 0.0;    /* initialize_J */
        for (int64_t k = 0; k < L; k++) s->Ls[p*L + k] = 0.0;   /* lambda.clear */
        if (s->band_n) for (int64_t k = 0; k < L; k++) s->band_n[p*L + k] = 0;
    }
    const int64_t nband = s->m.n_off_diag;
    int err = 0;
    if (nthreads <= 0) nthreads = orc_nthreads_max();
    for (int64_t r = 0; r < s->R / 2; r++) {
        const int64_t R = r;
        #pragma omp parallel num_threads(nthreads) reduction(|:err)
        {
            o_pair_buf b;
            int64_t cap = 256;
            o_buf_alloc(&b, cap, nband);
            int tid = 0, nth = 1;
#ifdef _OPENMP
            tid = omp_get_thread_num(); nth = omp_get_num_threads();
#endif
            for (int64_t oi = o_begin + tid; oi < o_end; oi += nth) {
                const int64_t o = list ? list[oi] : oi;
                const int64_t ar = s->antipod[r];
                const double weight_ang = s->m.weights[r];
                const double dshift_max = orc_get_dshift_max(s, o);
                int64_t n_r, n_ar_t;
                for (;;) {
                    n_r = orc_trace_ray(s, o, r, dshift_max, b.cap, NULL, b.cell_r, b.shift_r, b.dZ_r);
                    n_ar_t = orc_trace_ray(s, o, ar, dshift_max, b.cap, NULL, b.cell_ar, b.shift_ar, b.dZ_ar);
                    if (n_r > b.cap || n_ar_t > b.cap) {
                        cap = 2 * (n_r > n_ar_t ? n_r : n_ar_t);
                        o_buf_free(&b); o_buf_alloc(&b, cap, nband);
                        continue;
                    }
                    break;
                }
                if (n_r < 0 || n_ar_t < 0) { err = 1; continue; }
                for (int64_t i = 0; i < n_r; i++) b.lnotch_r[i] = 0;
                for (int64_t i = 0; i < n_ar_t; i++) b.lnotch_ar[i] = 0;
                /* the longer ray goes on the ar side (sim_radiation.tpp:164-171) */
                const int r_longer = (n_r > n_ar_t);
                const int64_t nA = r_longer ? n_r : n_ar_t, nB = r_longer ? n_ar_t : n_r;
                int64_t *cA = r_longer ? b.cell_r : b.cell_ar, *cB = r_longer ? b.cell_ar : b.cell_r;
                double *sA = r_longer ? b.shift_r : b.shift_ar, *sB = r_longer ? b.shift_ar : b.shift_r;
                double *zA = r_longer ? b.dZ_r : b.dZ_ar, *zB = r_longer ? b.dZ_ar : b.dZ_r;
                int64_t *lA = r_longer ? b.lnotch_r : b.lnotch_ar, *lB = r_longer ? b.lnotch_ar : b.lnotch_r;
                const int64_t ndep = nA + nB + 1;
                if (ndep > 1) {
                    int64_t lnotch_at_origin = 0;
                    for (int64_t f = 0; f < F; f++) {
                        double Lo, chio;
                        const double u = o_setup_and_solve(s, R, o, f, nA, cA, sA, zA, lA,
                                                           nB, cB, sB, zB, lB,
                                                           &lnotch_at_origin, &b, &Lo, &chio);
                        if (s->appears[f]) {
                            /* species l, its transition k (raypair_lambda.tpp:103-124) */
                            const int64_t l = s->corr_l[f], k = s->corr_k[f], z = s->corr_z[f];
                            const mgr_species_view *sv = &s->sp[l].v;
                            const int64_t g = s->sp[l].loff + k;
                            const double freq_line = sv->frequency[k];
                            const double weight = sv->qweights[z] * 2.0 * weight_ang;
                            const double factor = sv->A[k] * weight;
                            const double frs = s->nu[o*F + f];
                            const double prof = orc_profile(freq_line * o_profile_width(s, sv->inverse_mass, o),
                                                            frs - freq_line);
                            const double Lel = factor * (frs * prof * Lo / chio);
                            s->Ls[o*L + g] += Lel;
                            if (nband >= 2) o_update_bands(s, o, g, factor, freq_line, nA, nB, &b);
                        }
                        s->J[o*F + f] += 2.0 * weight_ang * u;
                        if (s->u) s->u[(R*s->N + o)*F + f] = u;   /* radiation.u[R][ind] */
                    }
                } else {
                    const int64_t bb = s->cell2boundary_nr[o];
                    for (int64_t f = 0; f < F; f++) {
                        const double I = s->I_bdy[(R*s->NB + bb)*F + f];
                        const double u = 0.5 * (I + I);
                        s->J[o*F + f] += 2.0 * weight_ang * u;
                        if (s->u) s->u[(R*s->N + o)*F + f] = u;
                    }
                }
            }
            o_buf_free(&b);
        }
    }
    return err ? -1 : 0;
}

/* ---- Level populations -------------------------------------------------- */

/* Simulation::compute_Jeff (src/Simulation/sim_lines.tpp:176-220): every Lambda*
 * element of (p, k), nr = p first, then the band elements in list order */
void orc_compute_Jeff(orc_state *s)
{
    const int64_t F = s->F, L = s->L, Q = s->Q;
    for (int64_t l = 0; l < s->S; l++) {
        const mgr_species_view *sv = &s->sp[l].v;
        const int64_t V = s->sp[l].V, loff = s->sp[l].loff;
        const double *pop = s->pop + s->N * s->sp[l].voff;
        #pragma omp parallel for schedule(static)
        for (int64_t p = 0; p < s->N; p++)
            for (int64_t k = 0; k < s->sp[l].L; k++) {
                const int64_t g = loff + k;
                double Jlin = 0.0;
                for (int64_t z = 0; z < Q; z++)
                    Jlin += sv->qweights[z] * s->J[p*F + s->nr_line[(p*L + g)*Q + z]];
                double diff = 0.0;
                diff += s->Ls[p*L + g] * pop[p*V + sv->irad[k]];
                if (s->band_n)
                    for (int64_t e = 0; e < s->band_n[p*L + g]; e++)
                        diff += s->band_Ls[(p*L + g)*s->band_cap + e]
                              * pop[s->band_nr[(p*L + g)*s->band_cap + e]*V + sv->irad[k]];
                s->Jlin[p*L + g] = Jlin;
                s->Jeff[p*L + g] = Jlin - HH_OVER_FOUR_PI * diff;
                s->Jdif[p*L + g] = HH_OVER_FOUR_PI * diff;
            }
    }
}

/* Dense LU with partial pivoting, row-major M [n][n], solves M x = y in place. */
static int o_lu_solve(int64_t n, double *M, double *y)
{
    for (int64_t c = 0; c < n; c++) {
        int64_t piv = c; double best = fabs(M[c*n + c]);
        for (int64_t r = c+1; r < n; r++) if (fabs(M[r*n + c]) > best) { best = fabs(M[r*n + c]); piv = r; }
        if (best == 0.0) return -1;
        if (piv != c) {
            for (int64_t j = 0; j < n; j++) { const double t = M[c*n+j]; M[c*n+j] = M[piv*n+j]; M[piv*n+j] = t; }
            const double t = y[c]; y[c] = y[piv]; y[piv] = t;
        }
        for (int64_t r = c+1; r < n; r++) {
            const double l = M[r*n + c] / M[c*n + c];
            M[r*n + c] = l;
            for (int64_t j = c+1; j < n; j++) M[r*n + j] -= l * M[c*n + j];
            y[r] -= l * y[c];
        }
    }
    for (int64_t r = n-1; r >= 0; r--) {
        double acc = y[r];
        for (int64_t j = r+1; j < n; j++) acc -= M[r*n + j] * y[j];
        y[r] = acc / M[r*n + r];
    }
    return 0;
}

/* The same elimination for one large system, rows updated in parallel. */
static int o_lu_solve_parallel(int64_t n, double *M, double *y)
{
    for (int64_t c = 0; c < n; c++) {
        int64_t piv = c; double best = fabs(M[c*n + c]);
        for (int64_t r = c+1; r < n; r++) if (fabs(M[r*n + c]) > best) { best = fabs(M[r*n + c]); piv = r; }
        if (best == 0.0) return -1;
        if (piv != c) {
            for (int64_t j = 0; j < n; j++) { const double t = M[c*n+j]; M[c*n+j] = M[piv*n+j]; M[piv*n+j] = t; }
            const double t = y[c]; y[c] = y[piv]; y[piv] = t;
        }
        #pragma omp parallel for schedule(static) if (n - c > 256)
        for (int64_t r = c+1; r < n; r++) {
            const double l = M[r*n + c] / M[c*n + c];
            M[r*n + c] = l;
            if (l != 0.0)
                for (int64_t j = c+1; j < n; j++) M[r*n + j] -= l * M[c*n + j];
            y[r] -= l * y[c];
        }
    }
    for (int64_t r = n-1; r >= 0; r--) {
        double acc = y[r];
        for (int64_t j = r+1; j < n; j++) acc -= M[r*n + j] * y[j];
        y[r] = acc / M[r*n + r];
    }
    return 0;
}

/* Largest N V for which the global rate equations are solved densely. */
#define O_GLOBAL_MAX 8192

/* update_using_statistical_equilibrium with band elements: the triplets of
 * lineProducingSpecies.tpp:270-388 over all cells, where an element (p, k, nr)
 * adds +v at (index(p, j_k), index(nr, i_k)) and -v at (index(nr, i_k),
 * index(nr, i_k)), v = -Ls opacity(p, k) (lineProducingSpecies.tpp:317-340),
 * solved as one system (history already shifted). */
static int o_stateq_global(orc_state *s)
{
    const mgr_model_view *m = &s->m;
    const int64_t V = s->V, L = s->L, NV = s->N * s->V;
    if (NV > O_GLOBAL_MAX) return -2;
    double *M = (double *)xcalloc((size_t)NV * NV, sizeof(double));
    double *y = (double *)xcalloc(NV, sizeof(double));
    int64_t ncol_max = 1;
    for (int64_t c = 0; c < m->ncolpar; c++) if (m->colpar_ncol[c] > ncol_max) ncol_max = m->colpar_ncol[c];
    double *Ce = (double *)xcalloc(ncol_max, sizeof(double));
    double *Cd = (double *)xcalloc(ncol_max, sizeof(double));
    for (int64_t p = 0; p < s->N; p++) {
        const double *pop = s->prev1 + p*V;   /* == population before the solve */
        const int64_t b = p*V;
        for (int64_t k = 0; k < L; k++) {
            const double Jeff = s->Jeff[p*L + k];
            const double v_IJ = m->A[k] + m->Bs[k] * Jeff;
            const double v_JI = m->Ba[k] * Jeff;
            const int64_t I = b + m->irad[k], J = b + m->jrad[k];
            if (m->jrad[k] != V-1) { M[J*NV + I] += v_IJ; M[J*NV + J] += -v_JI; }
            if (m->irad[k] != V-1) { M[I*NV + J] += v_JI; M[I*NV + I] += -v_IJ; }
        }
        for (int64_t k = 0; k < L; k++) {
            const double opacity = HH_OVER_FOUR_PI * (pop[m->jrad[k]] * m->Ba[k] - pop[m->irad[k]] * m->Bs[k]);
            const int64_t J = b + m->jrad[k];
            const int64_t pk = p*L + k;
            for (int64_t e = -1; e < s->band_n[pk]; e++) {
                const int64_t nr = e < 0 ? p : s->band_nr[pk*s->band_cap + e];
                const double Lel = e < 0 ? s->Ls[pk] : s->band_Ls[pk*s->band_cap + e];
                const double v_IJ = -Lel * opacity;
                const int64_t I = nr*V + m->irad[k];
                if (m->jrad[k] != V-1) M[J*NV + I] += v_IJ;
                if (m->irad[k] != V-1) M[I*NV + I] += -v_IJ;
            }
        }
        int64_t koff = 0;
        for (int64_t c = 0; c < m->ncolpar; c++) {
            double factor;
            orc_collision_rates(s, c, m->temperature[p], Ce, Cd, &factor);
            const double abn = m->abundance[p*m->nspecs + m->colpar_num[c]] * factor;
            for (int64_t k = 0; k < m->colpar_ncol[c]; k++) {
                const double v_IJ = Cd[k] * abn, v_JI = Ce[k] * abn;
                const int64_t I = b + m->icol[koff + k], J = b + m->jcol[koff + k];
                if (m->jcol[koff + k] != V-1) { M[J*NV + I] += v_IJ; M[J*NV + J] += -v_JI; }
                if (m->icol[koff + k] != V-1) { M[I*NV + J] += v_JI; M[I*NV + I] += -v_IJ; }
            }
            koff += m->colpar_ncol[c];
        }
        for (int64_t i = 0; i < V; i++) M[(b + V-1)*NV + b + i] += 1.0;
        y[b + V-1] = s->pop_tot[p];
    }
    const int rc = o_lu_solve_parallel(NV, M, y);
    if (rc == 0) memcpy(s->pop, y, NV * sizeof(double));
    free(M); free(y); free(Ce); free(Cd);
    return rc;
}

/* LineProducingSpecies::update_using_statistical_equilibrium
 * (lineProducingSpecies.tpp:249-495): history shift, per-cell assembly of the
 * same triplets (transpose convention, last row = conservation), one dense LU
 * per cell in place of the global SparseLU (block diagonal without band
 * elements); with band elements one global solve (o_stateq_global). */
int orc_update_using_statistical_equilibrium(orc_state *s)
{
    const int64_t L = s->L, NV = s->N * s->V;
    memcpy(s->prev3, s->prev2, NV * sizeof(double));
    memcpy(s->prev2, s->prev1, NV * sizeof(double));
    memcpy(s->prev1, s->pop,   NV * sizeof(double));
    if (s->band_n) return o_stateq_global(s);
    int err = 0;
    for (int64_t l = 0; l < s->S; l++) {   /* Lines::iteration_using_statistical_equilibrium */
        const mgr_species_view *m = &s->sp[l].v;
        const int64_t V = s->sp[l].V, loff = s->sp[l].loff;
        const double *prev1 = s->prev1 + s->N * s->sp[l].voff, *pop_tot = s->pop_tot + l * s->N;
        double *popl = s->pop + s->N * s->sp[l].voff;
        int64_t ncol_max = 1;
        for (int64_t c = 0; c < m->ncolpar; c++) if (m->colpar_ncol[c] > ncol_max) ncol_max = m->colpar_ncol[c];
        #pragma omp parallel reduction(|:err)
        {
            double *M = (double *)xcalloc(V*V, sizeof(double));
            double *y = (double *)xcalloc(V, sizeof(double));
            double *Ce = (double *)xcalloc(ncol_max, sizeof(double));
            double *Cd = (double *)xcalloc(ncol_max, sizeof(double));
            #pragma omp for schedule(static)
            for (int64_t p = 0; p < s->N; p++) {
                for (int64_t q = 0; q < V*V; q++) M[q] = 0.0;
                for (int64_t q = 0; q < V; q++) y[q] = 0.0;
                const double *pop = prev1 + p*V;   /* == population before the solve */
                /* radiative transitions */
                for (int64_t k = 0; k < m->nrad; k++) {
                    const double Jeff = s->Jeff[p*L + loff + k];
                    const double v_IJ = m->A[k] + m->Bs[k] * Jeff;
                    const double v_JI = m->Ba[k] * Jeff;
                    const int64_t I = m->irad[k], J = m->jrad[k];
                    if (J != V-1) { M[J*V + I] += v_IJ; M[J*V + J] += -v_JI; }
                    if (I != V-1) { M[I*V + J] += v_JI; M[I*V + I] += -v_IJ; }
                }
                /* approximated Lambda operator, kappa from the current populations */
                for (int64_t k = 0; k < m->nrad; k++) {
                    const int64_t I = m->irad[k], J = m->jrad[k];
                    const double opacity = HH_OVER_FOUR_PI * (pop[J] * m->Ba[k] - pop[I] * m->Bs[k]);
                    const double v_IJ = -s->Ls[p*L + loff + k] * opacity;
                    if (J != V-1) M[J*V + I] += v_IJ;
                    if (I != V-1) M[I*V + I] += -v_IJ;
                }
                /* collisional transitions */
                int64_t koff = 0;
                for (int64_t c = 0; c < m->ncolpar; c++) {
                    double factor;
                    const double tmp = s->m.temperature[p];
                    o_collision_rates(m, c, tmp, Ce, Cd, &factor);
                    const double abn = s->m.abundance[p*s->m.nspecs + m->colpar_num[c]] * factor;
                    for (int64_t k = 0; k < m->colpar_ncol[c]; k++) {
                        const double v_IJ = Cd[k] * abn, v_JI = Ce[k] * abn;
                        const int64_t I = m->icol[koff + k], J = m->jcol[koff + k];
                        if (J != V-1) { M[J*V + I] += v_IJ; M[J*V + J] += -v_JI; }
                        if (I != V-1) { M[I*V + J] += v_JI; M[I*V + I] += -v_IJ; }
                    }
                    koff += m->colpar_ncol[c];
                }
                for (int64_t i = 0; i < V; i++) M[(V-1)*V + i] += 1.0;
                y[V-1] = pop_tot[p];
                if (o_lu_solve(V, M, y) != 0) { err = 1; continue; }
                for (int64_t i = 0; i < V; i++) popl[p*V + i] = y[i];
            }
            free(M); free(y); free(Ce); free(Cd);
        }
    }
    return err ? -1 : 0;
}

/* LineProducingSpecies::check_for_convergence (lineProducingSpecies.tpp:120-171) */
void orc_check_for_convergence(orc_state *s, double pop_prec)
{
    s->fnc = s->rcm = s->rcmax = 0.0;
    for (int64_t l = 0; l < s->S; l++) {
        const int64_t V = s->sp[l].V;
        const double *pop = s->pop + s->N * s->sp[l].voff, *prev1 = s->prev1 + s->N * s->sp[l].voff;
        const double *pop_tot = s->pop_tot + l * s->N;
        const double weight = 1.0 / (double)(s->N * V);
        double fnc = 0.0, rcm = 0.0, rcmax = 0.0;
        #pragma omp parallel for reduction(+:fnc, rcm) reduction(max:rcmax)
        for (int64_t p = 0; p < s->N; p++) {
            const double min_pop = 1.0E-10 * pop_tot[p];
            for (int64_t i = 0; i < V; i++) {
                const int64_t ind = p*V + i;
                if (pop[ind] > min_pop) {
                    double relative_change = 2.0;
                    relative_change *= fabs(pop[ind] - prev1[ind]);
                    relative_change /= (pop[ind] + prev1[ind]);
                    if (relative_change > pop_prec) fnc += weight;
                    rcm += (weight * relative_change);
                    if (relative_change > rcmax) rcmax = relative_change;
                }
            }
        }
        s->sp[l].fnc = fnc; s->sp[l].rcm = rcm; s->sp[l].rcmax = rcmax;
        /* the ALI continues while any species is not converged (sim_lines.tpp:130-150):
         * the driver's figures are the largest over the species */
        if (fnc > s->fnc) s->fnc = fnc;
        if (rcm > s->rcm) s->rcm = rcm;
        if (rcmax > s->rcmax) s->rcmax = rcmax;
    }
}

/* LineProducingSpecies::update_using_Ng_acceleration (lineProducingSpecies.tpp:181-238),
 * for every species in turn (Lines::iteration_using_Ng_acceleration) */
void orc_update_using_Ng_acceleration(orc_state *s)
{
    for (int64_t l = 0; l < s->S; l++) {
        const int64_t NV = s->N * s->sp[l].V, b0 = s->N * s->sp[l].voff;
        double *pop = s->pop + b0, *prev1 = s->prev1 + b0, *prev2 = s->prev2 + b0, *prev3 = s->prev3 + b0;
        double A1 = 0.0, A2 = 0.0, B2 = 0.0, C1 = 0.0, C2 = 0.0;
        for (int64_t i = 0; i < NV; i++) {
            const double x = pop[i], x1 = prev1[i], x2 = prev2[i], x3 = prev3[i];
            const double Q1 = x - 2.0*x1 + x2;
            const double Q2 = x - x1 - x2 + x3;
            const double Q3 = x - x1;
            A1 += Q1*Q1; A2 += Q1*Q2; B2 += Q2*Q2; C1 += Q1*Q3; C2 += Q2*Q3;
        }
        const double B1 = A2;
        const double denominator = A1*B2 - A2*B1;
        if (denominator != 0.0) {
            const double a = (C1*B2 - C2*B1) / denominator;
            const double b = (C2*A1 - C1*A2) / denominator;
            for (int64_t i = 0; i < NV; i++) {
                const double x = pop[i];
                pop[i] = (1.0 - a - b)*x + a*prev1[i] + b*prev2[i];
                prev3[i] = prev2[i];
                prev2[i] = prev1[i];
                prev1[i] = x;
            }
        }
    }
}

/* Simulation::compute_level_populations (src/Simulation/sim_lines.tpp:61-164) */
int64_t orc_compute_level_populations(orc_state *s, int use_Ng, int64_t max_niterations,
                                      int criterion, int nthreads)
{
    if (!s->spec_set) return -1;
    int64_t iteration = 0, iteration_normal = 0;
    int some_not_converged = 1;
    while (some_not_converged && (iteration < max_niterations)) {
        iteration++;
        some_not_converged = 0;
        if (use_Ng && (iteration_normal == 4)) {
            orc_update_using_Ng_acceleration(s);
            orc_check_for_convergence(s, s->m.pop_prec);
            orc_set_emissivity_and_opacity(s);
            iteration_normal = 0;
        } else {
            if (orc_compute_radiation_field(s, nthreads, 0, s->N) != 0) return -1;
            orc_compute_Jeff(s);
            if (orc_update_using_statistical_equilibrium(s) != 0) return -1;
            orc_check_for_convergence(s, s->m.pop_prec);
            orc_set_emissivity_and_opacity(s);
            iteration_normal++;
        }
        if (criterion == MGR_CRIT_MAX) {
            if (s->rcmax >= s->m.pop_prec) some_not_converged = 1;
        } else {
            if (s->fnc > 0.005) some_not_converged = 1;
        }
    }
    return iteration;
}

/* ---- accessors ---------------------------------------------------------- */
void orc_get_frequencies(const orc_state *s, double *nu, int64_t *nr_line)
{
    if (nu) memcpy(nu, s->nu, s->N * s->F * sizeof(double));
    if (nr_line) memcpy(nr_line, s->nr_line, s->N * s->L * s->Q * sizeof(int64_t));
}
void orc_get_boundary_intensities(const orc_state *s, double *I_bdy)
{
    memcpy(I_bdy, s->I_bdy, s->NB * s->F * sizeof(double));
}
void orc_get_J(const orc_state *s, double *J) { memcpy(J, s->J, s->N * s->F * sizeof(double)); }
void orc_get_lambda(const orc_state *s, double *Ls) { memcpy(Ls, s->Ls, s->N * s->L * sizeof(double)); }
int64_t orc_lambda_band_capacity(const orc_state *s) { return s->band_cap; }
void orc_get_lambda_band(const orc_state *s, int64_t *count, int64_t *nr, double *Ls)
{
    if (!s->band_n) return;
    const size_t n = s->N * s->L;
    if (count) memcpy(count, s->band_n, n * sizeof(int64_t));
    if (nr) memcpy(nr, s->band_nr, n * s->band_cap * sizeof(int64_t));
    if (Ls) memcpy(Ls, s->band_Ls, n * s->band_cap * sizeof(double));
}
void orc_get_Jeff(const orc_state *s, double *Jlin, double *Jeff, double *Jdif)
{
    const size_t n = s->N * s->L * sizeof(double);
    if (Jlin) memcpy(Jlin, s->Jlin, n);
    if (Jeff) memcpy(Jeff, s->Jeff, n);
    if (Jdif) memcpy(Jdif, s->Jdif, n);
}
void orc_get_populations(const orc_state *s, double *pop) { memcpy(pop, s->pop, s->N * s->V * sizeof(double)); }
void orc_set_populations(orc_state *s, const double *pop)
{
    memcpy(s->pop, pop, s->N * s->V * sizeof(double));
    for (int64_t l = 0; l < s->S; l++)
        for (int64_t p = 0; p < s->N; p++)
            s->pop_tot[l*s->N + p] = s->m.abundance[p*s->m.nspecs + s->sp[l].v.num];
    orc_set_emissivity_and_opacity(s);
}
void orc_get_population_history(const orc_state *s, double *p1, double *p2, double *p3)
{
    const size_t n = s->N * s->V * sizeof(double);
    if (p1) memcpy(p1, s->prev1, n);
    if (p2) memcpy(p2, s->prev2, n);
    if (p3) memcpy(p3, s->prev3, n);
}
void orc_set_population_history(orc_state *s, const double *p1, const double *p2, const double *p3)
{
    const size_t n = s->N * s->V * sizeof(double);
    if (p1) memcpy(s->prev1, p1, n);
    if (p2) memcpy(s->prev2, p2, n);
    if (p3) memcpy(s->prev3, p3, n);
}
void orc_get_emissivity_opacity(const orc_state *s, double *emis, double *opac)
{
    const size_t n = s->N * s->L * sizeof(double);
    if (emis) memcpy(emis, s->emissivity, n);
    if (opac) memcpy(opac, s->opacity, n);
}
void orc_get_convergence(const orc_state *s, double *fnc, double *rcm, double *rcmax)
{
    *fnc = s->fnc; *rcm = s->rcm; *rcmax = s->rcmax;
}
int64_t orc_nlspecs(const orc_state *s) { return s->S; }
void orc_get_convergence_species(const orc_state *s, int64_t l, double *fnc, double *rcm, double *rcmax)
{
    *fnc = s->sp[l].fnc; *rcm = s->sp[l].rcm; *rcmax = s->sp[l].rcmax;
}
void orc_get_antipod(const orc_state *s, int64_t *antipod) { memcpy(antipod, s->antipod, s->R * sizeof(int64_t)); }
void orc_get_lines(const orc_state *s, double *line, int64_t *line_index)
{
    memcpy(line, s->line, s->L * sizeof(double));
    memcpy(line_index, s->line_index, s->L * sizeof(int64_t));
}
void orc_get_line_maps(const orc_state *s, int64_t *appears, int64_t *l, int64_t *k, int64_t *z)
{
    for (int64_t f = 0; f < s->F; f++) {
        appears[f] = s->appears[f]; l[f] = s->corr_l[f]; k[f] = s->corr_k[f]; z[f] = s->corr_z[f];
    }
}
