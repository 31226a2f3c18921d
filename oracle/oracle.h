/*
 * oracle.h -- CPU FP64 restatement of Magritte's radiation-field and
 * level-population path (the "cpu_ref" oracle).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker or
 * as the timed CPU baseline.  The product path (libmagritte_hip.so) never
 * links, loads or calls it.
 *
 * Every function follows the reference's operation order so that it can be
 * pinned against the reference's own known answers and closed forms; each
 * cites the reference file:line it restates.
 */
#ifndef MGR_ORACLE_H
#define MGR_ORACLE_H

#include <stdint.h>
#include "../include/magritte_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_state orc_state;

orc_state *orc_create(const mgr_model_view *m);
void       orc_destroy(orc_state *s);
int        orc_nthreads_max(void);

/* src/Simulation/sim_spectrum.tpp:13-196 */
int  orc_compute_spectral_discretisation(orc_state *s);
/* src/Simulation/sim_radiation.tpp:17-48; temperatures may be NULL (T_CMB) */
int  orc_compute_boundary_intensities(orc_state *s, const double *temperatures);
/* src/Simulation/sim_lines.tpp:10-16 -> lineProducingSpecies.tpp:82-115, lines.tpp:44-60 */
int  orc_compute_LTE_level_populations(orc_state *s);
/* src/Model/Lines/lines.tpp:44-60 */
void orc_set_emissivity_and_opacity(orc_state *s);

/* src/Simulation/simulation.tpp:15-37 */
double orc_get_dshift_max(const orc_state *s, int64_t o);
/* src/Model/Geometry/geometry.tpp:21-95 (CoMoving).  Writes at most cap points;
 * returns the number of points, or -1 if a ray found no neighbour (geometry.tpp:68-75). */
int64_t orc_trace_ray(const orc_state *s, int64_t o, int64_t r, double dshift_max,
                      int64_t cap, int64_t *crt, int64_t *cellNr, double *shift, double *dZ);
/* Geometry::trace_ray<Rest> (geometry.tpp:21-95, rest-frame shifts) */
int64_t orc_trace_ray_rest(const orc_state *s, int64_t o, int64_t r, double dshift_max,
                           int64_t cap, int64_t *crt, int64_t *cellNr, double *shift, double *dZ);
/* src/Simulation/sim_spectrum.tpp:206-311 */
int  orc_compute_spectral_discretisation_image(orc_state *s, double width);
/* RayPair::solve_for_image (raypair_solver.tpp:523-579): u over [0, last] */
void orc_solve_for_image(int64_t ndep, int64_t first, int64_t last, const double *term1,
                         const double *dtau, double I_bdy_0, double I_bdy_n, double *Su);
/* src/Simulation/sim_radiation.tpp:345-478 for ray r < nrays/2: I_m, I_p [ncells*nfreqs] */
int  orc_compute_image(orc_state *s, int64_t r, double *I_m, double *I_p);
/* Radiation::rescale_I_bdy (src/Model/Radiation/radiation.tpp:341-373): the
 * boundary intensity of boundary cell b (cell p) for reduced ray R at
 * freq_scaled, searched in cell p's frequency grid from notch 0. */
double orc_rescale_I_bdy(const orc_state *s, double freq_scaled, int64_t R, int64_t p, int64_t b);
/* Boundary::cell2boundary_nr (src/Model/Geometry/Boundary/boundary.cpp) */
int64_t orc_cell2boundary_nr(const orc_state *s, int64_t p);
/* src/Model/Geometry/geometry.tpp:498-543 */
int64_t orc_get_npoints_on_ray(const orc_state *s, int64_t o, int64_t r, double dshift_max);
/* src/Model/Geometry/geometry.tpp:433-465 */
double orc_get_doppler_shift(const orc_state *s, int64_t o, int64_t r, int64_t c, int rest_frame);

/* RayPair::solve (src/Simulation/Raypair/raypair_solver.tpp:200-420) on given
 * term1 / dtau; first, last, n_ar as in the reference.  Outputs Su[ndep],
 * L_diag[ndep] and, for n_off_diag > 0, L_upper/L_lower [n_off_diag][ndep]. */
void orc_raypair_solve(int64_t ndep, int64_t n_ar, int64_t first, int64_t last,
                       int64_t n_off_diag, const double *term1, const double *dtau,
                       double I_bdy_0, double I_bdy_n,
                       double *Su, double *L_diag, double *L_upper, double *L_lower,
                       double *A_out, double *C_out);

/* src/Simulation/sim_radiation.tpp:56-338 for origins [o_begin, o_end) of
 * every reduced ray; nthreads <= 0 uses all OpenMP threads.  J and Lambda of
 * the covered origins are reset first. */
/* Radiation::u with scattering: u[(R*N + p)*F + f] of every ray pair R < nrays/2 */
void orc_set_scattering(orc_state *s, int use_scattering);
void orc_get_u(const orc_state *s, double *u);
int  orc_compute_radiation_field(orc_state *s, int nthreads, int64_t o_begin, int64_t o_end);
/* the same for an explicit list of origins (bounded CPU-baseline samples) */
int  orc_compute_radiation_field_list(orc_state *s, int nthreads, const int64_t *origins, int64_t n);
/* src/Simulation/sim_lines.tpp:176-220 */
void orc_compute_Jeff(orc_state *s);
/* lineProducingSpecies.tpp:249-495 (per-cell dense LU; with band elements one
 * dense LU over all N V unknowns, at most 8192; see oracle.c) */
int  orc_update_using_statistical_equilibrium(orc_state *s);
/* lineProducingSpecies.tpp:120-171 */
void orc_check_for_convergence(orc_state *s, double pop_prec);
/* lineProducingSpecies.tpp:181-238 */
void orc_update_using_Ng_acceleration(orc_state *s);
/* src/Simulation/sim_lines.tpp:61-164 */
int64_t orc_compute_level_populations(orc_state *s, int use_Ng, int64_t max_niterations,
                                      int criterion, int nthreads);

/* accessors: copy state arrays out / in.  Populations (and their history)
 * hold the species one after the other, each as index(p, i) = i + p*nlev;
 * per-line arrays (Ls, Jlin, Jeff, Jdif, emissivity, opacity) are
 * [ncells][nlines] with the species' lines one after the other. */
void orc_get_frequencies(const orc_state *s, double *nu, int64_t *nr_line);
void orc_get_boundary_intensities(const orc_state *s, double *I_bdy /* [nboundary*F] for R=0 */);
void orc_get_J(const orc_state *s, double *J);
void orc_get_lambda(const orc_state *s, double *Ls);   /* element nr = p of every (p, k) */
/* n_off_diag >= 2: the other elements of (p, k), count[p*L+k] of them at
 * (p*L+k)*capacity in nr / Ls (Lambda::add_element order, lambda.tpp:130-143) */
int64_t orc_lambda_band_capacity(const orc_state *s);
void orc_get_lambda_band(const orc_state *s, int64_t *count, int64_t *nr, double *Ls);
void orc_get_Jeff(const orc_state *s, double *Jlin, double *Jeff, double *Jdif);
void orc_get_populations(const orc_state *s, double *pop);
void orc_set_populations(orc_state *s, const double *pop);
void orc_get_population_history(const orc_state *s, double *p1, double *p2, double *p3);
void orc_set_population_history(orc_state *s, const double *p1, const double *p2, const double *p3);
void orc_get_emissivity_opacity(const orc_state *s, double *emis, double *opac);
void orc_get_convergence(const orc_state *s, double *fnc, double *rcm, double *rcmax);  /* largest over species */
int64_t orc_nlspecs(const orc_state *s);
void orc_get_convergence_species(const orc_state *s, int64_t l, double *fnc, double *rcm, double *rcmax);
void orc_get_antipod(const orc_state *s, int64_t *antipod);
void orc_get_lines(const orc_state *s, double *line, int64_t *line_index);
/* appears_in_line_integral, corresponding_{l,k,z} (length nfreqs) */
void orc_get_line_maps(const orc_state *s, int64_t *appears, int64_t *l, int64_t *k, int64_t *z);

/* Closed-form helpers used by tests */
double orc_planck(double temperature, double freq);             /* src/Functions/planck.tpp:15-24 */
double orc_profile(double width, double freq_diff);             /* thermodynamics.tpp:16-29 */
double orc_profile_width(const orc_state *s, int64_t p);        /* thermodynamics.tpp:62-69 */
void   orc_get_eta_and_chi(const orc_state *s, double freq_scaled, int64_t p,
                           int64_t *lnotch, double *eta, double *chi); /* simulation.tpp:385-481 */
/* src/Model/Lines/LineProducingSpecies/Linedata/CollisionPartner/collisionPartner.tpp:15-72 */
void   orc_collision_rates(const orc_state *s, int64_t c, double temperature,
                           double *Ce_intpld, double *Cd_intpld, double *abundance_factor);

#ifdef __cplusplus
}
#endif
#endif
