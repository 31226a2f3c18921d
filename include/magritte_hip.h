/*
 * magritte_hip.h -- C-ABI of the MI355X (gfx950) radiation-field and
 * level-population library (libmagritte_hip.so).
 *
 * Plain C: fixed-width integers, plain pointers and sizes, no C++ or torch
 * types.  Every entry point returns an int status (MGR_OK = 0, negative on
 * failure) and never throws; the message of the last failure on a context is
 * returned by mgr_last_error().
 *
 * Each entry point names the reference (UCL/Magritte) interface it replaces,
 * as path:line inside the reference tree.  The model is described by the
 * io-layout arrays the reference reads in Model::read (src/Model/model.cpp:54-86):
 * units are SI, velocities in units of c, vturb2 in (v/c)^2.
 *
 * One line-producing species per context (the reference loops over
 * lines.lineProducingSpecies; the configurations served here all have one).
 */
#ifndef MAGRITTE_HIP_H
#define MAGRITTE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes                                                              */
/* ------------------------------------------------------------------------ */
#define MGR_OK                 0
#define MGR_ERR_ARGUMENT      -1   /* bad argument / precondition (reference returns -1) */
#define MGR_ERR_DEVICE        -2   /* HIP runtime error                                 */
#define MGR_ERR_NOMEM         -3   /* device allocation failed                          */
#define MGR_ERR_TRACE         -4   /* a ray found no forward neighbour inside the mesh  */
#define MGR_ERR_NUMERIC       -5   /* singular rate matrix                              */
#define MGR_ERR_STATE         -6   /* called out of order (e.g. no spectral discretisation) */
#define MGR_ERR_UNSUPPORTED   -7   /* feature outside this library's scope              */

/* Precision of the radiation-field solve (opacity, emissivity, source function,
 * Feautrier elimination).  J is accumulated over rays in FP64 in both modes. */
#define MGR_FP32 0
#define MGR_FP64 1

/* Convergence criteria of the ALI driver. */
#define MGR_CRIT_FRACTION 0  /* fraction of levels with rel. change > pop_prec must be <= 0.005
                                (src/Simulation/sim_lines.tpp:140)                          */
#define MGR_CRIT_MAX      1  /* max relative change over levels with n > 1e-10 n_tot < pop_prec
                                (magritte/run/levelpops.py:11-13,44)                        */

/* ------------------------------------------------------------------------ */
/* Model view: what Model::read loads (src/Model/model.cpp:54-86)            */
/* ------------------------------------------------------------------------ */

/* One further line-producing species, Lines/LineProducingSpecies_<l> for
 * l >= 1 (LineProducingSpecies::read, lineProducingSpecies.cpp:21-134): the
 * same fields, with the same meaning, as species 0's in mgr_model_view. */
typedef struct mgr_species_view {
    int64_t        num;
    double         inverse_mass;
    int64_t        nlev;
    int64_t        nrad;
    const double  *energy;
    const double  *weight;
    const int64_t *irad;
    const int64_t *jrad;
    const double  *A;
    const double  *Bs;
    const double  *Ba;
    const double  *frequency;
    int64_t        ncolpar;
    const int64_t *colpar_num;
    const char    *colpar_orth_or_para;
    const int64_t *colpar_ncol;
    const int64_t *colpar_ntmp;
    const int64_t *icol;
    const int64_t *jcol;
    const double  *tmp;
    const double  *Cd;
    const double  *Ce;
    const double  *roots;        /* [nquads] this species' quadrature            */
    const double  *qweights;     /* [nquads]                                      */
} mgr_species_view;

typedef struct mgr_model_view {
    /* Parameters (src/Model/parameters.hpp:18-96) */
    int64_t ncells;
    int64_t nrays;
    int64_t nboundary;
    int64_t nspecs;
    int64_t nquads;
    double  pop_prec;            /* required precision for populations           */
    double  max_width_fraction;  /* Parameters::max_width_fraction, default 0.5  */
    int64_t n_off_diag;          /* ALO bandwidth: 0, or 1 (the same diagonal operator) */

    /* Geometry/Cells (src/Model/Geometry/Cells/cells.cpp:20-93) */
    const double  *position;     /* [ncells][3] m                                 */
    const double  *velocity;     /* [ncells][3] units of c                        */
    const int64_t *n_neighbors;  /* [ncells]                                      */
    const int64_t *neighbors;    /* [sum(n_neighbors)] flattened, in list order   */

    /* Geometry/Rays (src/Model/Geometry/Rays/rays.cpp:22-52) */
    const double  *rays;         /* [nrays][3] unit vectors                       */
    const double  *weights;      /* [nrays]                                       */

    /* Geometry/Boundary (src/Model/Geometry/Boundary/boundary.cpp:19-67) */
    const int64_t *boundary2cell_nr;      /* [nboundary]                          */

    /* Thermodynamics (src/Model/Thermodynamics) */
    const double  *temperature;  /* [ncells] gas temperature, K                   */
    const double  *vturb2;       /* [ncells] (v_turb/c)^2                         */

    /* Chemistry/Species (src/Model/Chemistry/Species/species.cpp:19-56) */
    const double  *abundance;    /* [ncells][nspecs] number densities, m^-3       */

    /* Lines/LineProducingSpecies_0/Linedata (Linedata/linedata.cpp:25-88) */
    int64_t        num;          /* species index of the line producer            */
    double         inverse_mass; /* 1 / (mass in amu)                             */
    int64_t        nlev;
    int64_t        nrad;
    const double  *energy;       /* [nlev] J                                      */
    const double  *weight;       /* [nlev] statistical weights                    */
    const int64_t *irad;         /* [nrad] upper level                            */
    const int64_t *jrad;         /* [nrad] lower level                            */
    const double  *A;            /* [nrad] Einstein A, s^-1                       */
    const double  *Bs;           /* [nrad] stimulated emission coefficient        */
    const double  *Ba;           /* [nrad] absorption coefficient                 */
    const double  *frequency;    /* [nrad] Hz                                     */

    /* .../Linedata/CollisionPartner_<c> (CollisionPartner/collisionPartner.cpp:20-60),
     * concatenated over partners c = 0..ncolpar-1                                */
    int64_t        ncolpar;
    const int64_t *colpar_num;   /* [ncolpar] species index of the partner        */
    const char    *colpar_orth_or_para; /* [ncolpar] 'n', 'o' or 'p'              */
    const int64_t *colpar_ncol;  /* [ncolpar]                                     */
    const int64_t *colpar_ntmp;  /* [ncolpar]                                     */
    const int64_t *icol;         /* [sum ncol] upper level                        */
    const int64_t *jcol;         /* [sum ncol] lower level                        */
    const double  *tmp;          /* [sum ntmp] table temperatures, K              */
    const double  *Cd;           /* [sum ntmp*ncol], partner-major, then [t][k], m^3 s^-1 */
    const double  *Ce;           /* same layout                                   */

    /* Lines/LineProducingSpecies_0/Quadrature (Quadrature/quadrature.cpp:19-40) */
    const double  *roots;        /* [nquads]                                      */
    const double  *qweights;     /* [nquads]                                      */

    /* Parameters::nlspecs (parameters.hpp:18-96): the number of line-producing
     * species; species 0 is described above and species 1 .. nlspecs-1 by
     * more_species[0 .. nlspecs-2].  0 reads as 1.  The species' lines are
     * merged into one sorted line list (Lines::setup, lines.cpp:43-87); level
     * populations are then held species after species, each in its own
     * layout index(p, i) = i + p*nlev (lineProducingSpecies.tpp:24-30).      */
    int64_t                  nlspecs;
    const mgr_species_view  *more_species;
} mgr_model_view;

/* Convergence statistics of one level-population update
 * (LineProducingSpecies::check_for_convergence,
 *  src/Model/Lines/LineProducingSpecies/lineProducingSpecies.tpp:120-171).      */
typedef struct mgr_conv {
    double fraction_not_converged;
    double relative_change_mean;
    double relative_change_max;   /* computed here; the reference leaves it at 0 */
} mgr_conv;

typedef struct mgr_ctx mgr_ctx;   /* opaque: device-resident model + state       */

/* ------------------------------------------------------------------------ */
/* Context                                                                   */
/* ------------------------------------------------------------------------ */

/* Uploads the model to device `device` (SoA layout) and builds the derived
 * static tables.  Replaces Model::read (src/Model/model.cpp:54-86) plus the
 * antipodal-ray set-up Rays::setup_antipodal_rays (Rays/rays.cpp:90-103).     */
int  mgr_create(const mgr_model_view *model, int device, mgr_ctx **out);
/* A batch of independent models in one context (no counterpart in the
 * reference, which solves one model per process).  The models share the
 * species, line data, collision tables, quadrature and parameters; geometry,
 * rays, thermodynamics and abundances differ.  Cells, boundary cells and all
 * per-cell arrays of the getters and setters are concatenated in model order
 * (model m owns cells [cell_offsets[m], cell_offsets[m+1])).  Every kernel
 * covers the whole batch in one launch; mgr_compute_level_populations
 * iterates each model until it meets the criterion, freezing converged ones,
 * and Ng acceleration and convergence statistics are per model, so each
 * model follows the iteration it would follow alone.                        */
int  mgr_create_batch(const mgr_model_view *models, int64_t nmodels, int device, mgr_ctx **out);
int  mgr_batch_info(mgr_ctx *ctx, int64_t *nmodels, int64_t *cell_offsets /* [nmodels+1] or NULL */);
/* Per model, after mgr_compute_level_populations: the iteration at which it
 * converged (or the last one) and its last convergence statistics.          */
int  mgr_batch_convergence(mgr_ctx *ctx, int64_t *iterations, mgr_conv *conv);
/* The line-producing species of the context (Parameters::nlspecs) and each
 * one's nlev and nrad (arrays of nlspecs, may be NULL).                      */
int  mgr_species_info(mgr_ctx *ctx, int64_t *nlspecs, int64_t *nlev, int64_t *nrad);
/* Species l's check_for_convergence figures of the last level-population
 * update (LineProducingSpecies::fraction_not_converged, relative_change_mean,
 * relative_change_max; lineProducingSpecies.tpp:120-171); the mgr_conv of the
 * driver entry points holds the largest over the species.                   */
int  mgr_species_convergence(mgr_ctx *ctx, int64_t l, mgr_conv *conv);
void mgr_destroy(mgr_ctx *ctx);
const char *mgr_last_error(const mgr_ctx *ctx);   /* ctx may be NULL (create errors) */
int  mgr_device_count(int *count);

/* ------------------------------------------------------------------------ */
/* Model set-up (once per model)                                             */
/* ------------------------------------------------------------------------ */

/* Simulation::compute_spectral_discretisation (src/Simulation/sim_spectrum.tpp:13-196) */
int mgr_compute_spectral_discretisation(mgr_ctx *ctx);

/* Simulation::compute_spectral_discretisation_image (src/Simulation/sim_spectrum.tpp:206-311):
 * the same sorted grid with nu = line + line*width*root in every cell (width
 * a fraction of the line frequency).  Level populations then refuse to run
 * until compute_spectral_discretisation is called again (sim_lines.tpp:67-71). */
int mgr_compute_spectral_discretisation_image(mgr_ctx *ctx, double width);

/* Simulation::compute_boundary_intensities (src/Simulation/sim_radiation.tpp:17-48).
 * temperatures: [nboundary] K, or NULL for the CMB (T_CMB = 2.72548 K).      */
int mgr_compute_boundary_intensities(mgr_ctx *ctx, const double *temperatures);

/* Simulation::compute_LTE_level_populations (src/Simulation/sim_lines.tpp:10-16) */
int mgr_compute_LTE_level_populations(mgr_ctx *ctx);

/* Geometry::trace_ray<CoMoving> for every (origin, ray), cached on the device
 * (src/Model/Geometry/geometry.tpp:21-207,328-465).  The reference re-traces
 * inside every compute_radiation_field; traces here are computed once.       */
int mgr_trace(mgr_ctx *ctx);
/* Options of the next mgr_trace (no counterpart in the reference; traces are
 * bit-identical under every setting).  record_steps: how many cells of each
 * walk the counting pass records for the filling pass to replay instead of
 * scanning the neighbour lists again (-1: the default, 256; 0: none; else a
 * multiple of 4 up to 4096; the buffer is also held to a quarter of the free
 * memory).  candidate_masks: scan only the neighbours ahead of the current
 * cell along the ray (on by default; lists longer than 63 entries are always
 * scanned whole).                                                             */
int mgr_set_trace_options(mgr_ctx *ctx, int64_t record_steps, int candidate_masks);

/* ------------------------------------------------------------------------ */
/* Hot path                                                                  */
/* ------------------------------------------------------------------------ */

/* Simulation::compute_radiation_field (src/Simulation/sim_radiation.tpp:56-338):
 * J(p,f) = sum_r 2 w_r u_r(p,f) and the Lambda* diagonal, n_off_diag = 0.     */
int mgr_compute_radiation_field(mgr_ctx *ctx, int precision);

/* Parameters::use_scattering: Radiation::u (src/Model/Radiation/radiation.cpp:117-126).
 * With scattering on, every radiation field also keeps each ray pair's u at
 * the origin (sim_radiation.tpp:226-237, 258-262), [nrays/2][ncells*nfreqs]
 * doubles at (r*ncells + p)*nfreqs + f, read with mgr_get_u.                 */
int mgr_set_scattering(mgr_ctx *ctx, int use_scattering);
int mgr_get_u(mgr_ctx *ctx, double *u);

/* Simulation::compute_Jeff (src/Simulation/sim_lines.tpp:176-220)             */
int mgr_compute_Jeff(mgr_ctx *ctx);

/* Simulation::compute_level_populations_from_stateq (src/Simulation/sim_lines.tpp:24-30):
 * per-cell statistical equilibrium, convergence check, emissivity/opacity.    */
int mgr_compute_level_populations_from_stateq(mgr_ctx *ctx, mgr_conv *conv);

/* Lines::iteration_using_Ng_acceleration (src/Model/Lines/lines.cpp:131-149)  */
int mgr_iteration_using_Ng_acceleration(mgr_ctx *ctx, mgr_conv *conv);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  The Ng step and convergence statistics of
 * mgr_iteration_using_Ng_acceleration (Lines::iteration_using_Ng_acceleration,
 * lines.cpp:131-149, without the emissivity/opacity update) on caller-owned
 * device arrays of one model on the current device: population, prev1..3
 * [ncells*nlev] (index i + p*nlev) are updated in place as that entry point
 * updates the context's, population_tot [ncells] is read, and conv (a host
 * pointer, or NULL) receives the statistics after the stream is synchronised.
 * stream is a hipStream_t (NULL: the null stream).  It calls no allocation
 * routine: its scratch is a 196,640-byte static device array per GPU, which
 * limits ncells to 4,194,304 (and nlev to 4096); larger sizes return
 * MGR_ERR_ARGUMENT with the reason in mgr_last_error(NULL).  All calls on one
 * GPU share that array, so calls that overlap on one GPU (other streams,
 * other host threads) race.                                                   */
int mgr_ng_acceleration_noalloc(int64_t ncells, int64_t nlev, double pop_prec, double *population,
                                double *prev1, double *prev2, double *prev3, const double *population_tot,
                                void *stream, mgr_conv *conv);

/* Simulation::compute_level_populations(io, use_Ng, max_niterations)
 * (src/Simulation/sim_lines.tpp:61-164): the ALI loop, one hipGraph replay per
 * iteration.  iterations_out receives the iteration count (the reference's
 * return value).                                                              */
int mgr_compute_level_populations(mgr_ctx *ctx, int use_Ng_acceleration,
                                  int64_t max_niterations, int criterion,
                                  int precision, int64_t *iterations_out);

/* Builds what the first radiation field of a precision would build once per
 * model -- the trace, the out-of-window masks, the far-end intensity cache,
 * the launch-shape tuning and the ALI iteration graphs -- so that later calls
 * time the iterations alone (no counterpart in the reference).               */
int mgr_prepare(mgr_ctx *ctx, int precision);

/* One iteration of that loop (sim_lines.tpp:101-132), replayed from the same
 * hipGraphs: ng != 0 runs Lines::iteration_using_Ng_acceleration, otherwise
 * compute_radiation_field + compute_Jeff + iteration_using_statistical_equilibrium.
 * Lets a host driver keep the reference's per-iteration bookkeeping.          */
int mgr_ali_iteration(mgr_ctx *ctx, int ng, int precision, mgr_conv *conv);

/* ------------------------------------------------------------------------ */
/* Data movement (the reference keeps these as Model members)                */
/* ------------------------------------------------------------------------ */

/* LineProducingSpecies::population [ncells*nlev], index i + p*nlev.          */
int mgr_set_populations(mgr_ctx *ctx, const double *population);
int mgr_get_populations(mgr_ctx *ctx, double *population);
/* population_prev1..3 (Ng history), each [ncells*nlev]; any pointer may be NULL. */
int mgr_get_population_history(mgr_ctx *ctx, double *prev1, double *prev2, double *prev3);
int mgr_set_population_history(mgr_ctx *ctx, const double *prev1, const double *prev2,
                               const double *prev3);
/* Radiation::J [ncells*nfreqs], index f + p*nfreqs (radiation.tpp:11-15).    */
int mgr_get_J(mgr_ctx *ctx, double *J);
/* Lambda::Ls element nr = p of every (p,k), [ncells*nrad]: the whole operator
 * for n_off_diag <= 1.                                                        */
int mgr_get_lambda(mgr_ctx *ctx, double *Ls);
/* The other elements of Lambda (n_off_diag >= 2, RayPair::update_Lambda's
 * bands, raypair_lambda.tpp:129-146; Lambda::get_nr / get_Ls,
 * lambda.tpp:82-123): per cell `slots` = (nrays/2) 2 (n_off_diag-1) slots,
 * slot (r 2 + side) (n_off_diag-1) + m-1 for ray pair r and band m, side 0 on
 * the pair's longer ray (point n_ar - m), 1 on the shorter (n_ar + m).
 * cells [ncells*slots]: the slot's emitting cell nr, -1 where the element
 * does not exist; Ls [ncells*slots*nrad].  Elements of one (p, k) that name
 * the same cell add (Lambda::add_element, lambda.tpp:130-143).              */
int mgr_lambda_band_slots(mgr_ctx *ctx, int64_t *slots);
int mgr_get_lambda_band(mgr_ctx *ctx, int32_t *cells, double *Ls);
/* With band elements the rate equations (lineProducingSpecies.tpp:249-495,
 * one SparseLU over all cells in the reference) are solved by block-Jacobi
 * sweeps.  Of the last solve: the sweeps that moved a population (at most 16)
 * and, per sweep, the largest move over the cell's total population
 * (changes [16], may be NULL).  No counterpart in the reference.          */
int mgr_band_sweeps(mgr_ctx *ctx, int *sweeps, double *changes);
/* Jlin, Jeff, Jdif [ncells*nrad]; any pointer may be NULL.                    */
int mgr_get_Jeff(mgr_ctx *ctx, double *Jlin, double *Jeff, double *Jdif);
/* Frequencies::nu [ncells*nfreqs] and LineProducingSpecies::nr_line
 * [ncells*nrad*nquads].                                                       */
int mgr_get_frequencies(mgr_ctx *ctx, double *nu, int64_t *nr_line);
/* Radiation::I_bdy for one reduced ray: [nboundary*nfreqs] (all rays share it). */
int mgr_get_boundary_intensities(mgr_ctx *ctx, double *I_bdy);
/* Lines::emissivity / opacity [ncells*nrad] (lines.tpp:44-60).               */
int mgr_get_emissivity_opacity(mgr_ctx *ctx, double *emissivity, double *opacity);

/* Traced projected cells of one ray, expanded with the interpolation points
 * of Geometry::set_data (geometry.tpp:100-207) exactly as the solver consumes
 * them: cellNr, shift, dZ per point.  n_out receives the number of points;
 * at most `capacity` are written.                                             */
int mgr_get_ray(mgr_ctx *ctx, int64_t origin, int64_t ray, int64_t capacity,
                int64_t *cellNr, double *shift, double *dZ, int64_t *n_out);
/* Simulation::get_npoints_on_rays<CoMoving> (simulation.tpp:540-552):
 * npoints[(r)*ncells + o] = n_ar + n_r + 1 for r < nrays/2.                   */
int mgr_get_npoints_on_rays(mgr_ctx *ctx, int64_t *npoints);

/* Simulation::get_npoints_on_rays<Rest> (simulation.tpp:540-552) with
 * Geometry::get_doppler_shift<Rest> (geometry.tpp:448-463): the same counts in
 * the rest frame, for a single model.                                         */
int mgr_get_npoints_on_rays_rest(mgr_ctx *ctx, int64_t *npoints);

/* Simulation::compute_and_write_image (src/Simulation/sim_radiation.tpp:345-478)
 * with RayPair::solve_for_image (raypair_solver.tpp:523-579), for
 * 0 <= ray < nrays/2 of a single model, after
 * mgr_compute_spectral_discretisation_image and mgr_compute_boundary_intensities:
 * the intensities leaving each origin's rest-frame ray pair,
 * I_m[o*nfreqs + f] at the antipode's end (RayPair::get_Im_at_front) and
 * I_p[o*nfreqs + f] at the ray's end (RayPair::get_Ip_at_end).  Either
 * pointer may be NULL.  The image-plane coordinates (Image::set_coordinates,
 * image.cpp:93-132) are host arithmetic and stay with the caller.            */
int mgr_compute_image(mgr_ctx *ctx, int64_t ray, int precision, double *I_m, double *I_p);

/* Launch shape of the radiation-field kernel chosen by its autotuner (no
 * counterpart in the reference): waves of 64 frequency bins per workgroup and
 * traced steps staged per chunk, 0 before the first radiation field of that
 * precision; with capacity > 0 also the shapes tried by the last tuning and
 * their times on the tuning slice (ms).                                       */
int mgr_radiation_field_shape(mgr_ctx *ctx, int precision, int *waves_per_group, int *steps_per_chunk,
                              int capacity, int *n_tried, int *tried_waves, int *tried_steps,
                              float *tried_ms);
/* Forces that launch shape for later radiation fields of that precision
 * instead of tuning (no counterpart in the reference): 1..4 waves per
 * workgroup, 2..(the model's LDS limit, at most 64) steps per chunk; both 0
 * hands the choice back to the autotuner.  J and Lambda* do not depend on
 * the shape.                                                                  */
int mgr_set_radiation_field_shape(mgr_ctx *ctx, int precision, int waves_per_group, int steps_per_chunk);

/* Out-of-window skip of the radiation-field kernels (no counterpart in the
 * reference; on by default): steps at which no bin of a 64-bin wave comes
 * within 10.5 line widths of any line (FP32; the profiles are exactly 0 there)
 * or 11 (1 + 1e-6) widths (FP64; get_eta_and_chi's own window holds no line)
 * are eliminated with eta = 0 and chi at its floor without evaluating the
 * profiles; results are bit-identical either way.  Off evaluates every point. */
int mgr_set_window_skip(mgr_ctx *ctx, int enable);
/* Far-end intensity cache (no counterpart in the reference; on by default):
 * the rescaled boundary intensity entering every (origin, ray) and bin,
 * Radiation::rescale_I_bdy at the last point's shift, kept in the kernel's
 * precision once per trace, frequency grid and boundary intensities (4 B per
 * ray and bin for FP32, 8 B for FP64, when that fits three quarters of the
 * free device memory) instead of searched for in every radiation field.
 * Results are bit-identical either way; off frees it.                        */
int mgr_set_far_cache(mgr_ctx *ctx, int enable);
/* Interpolated points x 64-bin chunks of the cached trace, and how many of
 * them the skip leaves out (chunks beyond the fifth are never skipped).      */
int mgr_window_skip_stats(mgr_ctx *ctx, int64_t *point_chunks, int64_t *skipped);

/* Kernel time of the last mgr_compute_radiation_field, measured with HIP
 * events on the context's stream (ms).                                        */
int mgr_last_radiation_field_ms(mgr_ctx *ctx, double *ms);
/* Name of the radiation-field kernel that launch dispatched (the FP32 kernel
 * has a one-candidate-line instantiation), NUL-terminated in at most
 * `capacity` bytes; "" before the first radiation field.  No counterpart in
 * the reference.                                                              */
int mgr_last_radiation_field_kernel(mgr_ctx *ctx, char *name, int capacity);
/* Sum over the cached traces of the expanded points of every ray pair
 * (sum of ndep) and the number of ray pairs with ndep > 1.                    */
int mgr_trace_stats(mgr_ctx *ctx, int64_t *sum_ndep, int64_t *n_pairs_solved);

/* ---- learned warm start --------------------------------------------------
 * No counterpart in the reference, whose iteration starts from LTE
 * (Simulation::compute_LTE_level_populations, sim_lines.tpp:10-16).  A
 * multilayer perceptron 8 -> 128 -> 128 -> 7 (GELU; bfloat16 matrix-core
 * products accumulated in FP32; FP32 master weights and Adam) predicts
 * per-cell log10 level-population fractions; mgr_warm_start writes the
 * resulting populations into a context, after which the ALI loop runs
 * unchanged.  Parameters are packed as W1 [8][128], b1 [128], W2 [128][128],
 * b2 [128], W3 [128][7], b3 [7] (weights [in][out], row-major FP32).        */
#define MGR_MLP_IN     8
#define MGR_MLP_HIDDEN 128
#define MGR_MLP_OUT    7
typedef struct mgr_mlp mgr_mlp;   /* opaque: device-resident network          */

int64_t     mgr_mlp_num_params(void);
int         mgr_mlp_create(int device, uint64_t seed, mgr_mlp **out);
void        mgr_mlp_destroy(mgr_mlp *mlp);
const char *mgr_mlp_last_error(const mgr_mlp *mlp);
int         mgr_mlp_set_params(mgr_mlp *mlp, const float *params);   /* resets Adam */
int         mgr_mlp_get_params(mgr_mlp *mlp, float *params);
/* Inputs are normalised as (x - x_mean)/x_std, outputs de-normalised as
 * y' y_std + y_mean; the defaults are 0 and 1.                               */
int mgr_mlp_set_normalisation(mgr_mlp *mlp, const float *x_mean, const float *x_std,
                              const float *y_mean, const float *y_std);
/* Y [n][7] = network(X [n][8]); host buffers.                               */
int mgr_mlp_forward(mgr_mlp *mlp, const float *X, int64_t n, float *Y);
/* Mean squared error over rows and outputs of the normalised targets T [n][7]
 * and its gradient, packed like the parameters (grad may be NULL).           */
int mgr_mlp_loss_grad(mgr_mlp *mlp, const float *X, const float *T, int64_t n, float *loss,
                      float *grad);
/* Adam (beta 0.9 / 0.999, eps 1e-8) over shuffled minibatches on the device;
 * loss_per_epoch [epochs] may be NULL.                                        */
int mgr_mlp_train(mgr_mlp *mlp, const float *X, const float *T, int64_t n, int64_t epochs,
                  int64_t batch, float lr, uint64_t seed, float *loss_per_epoch);
/* Per-cell features [ncells][8]: log10 n(collision partner), log10 T,
 * log10 n(species), log10 of the largest neighbour velocity gradient (s^-1),
 * log10 Jlin_k / B_nu_k(T) for transitions k = 0..3 after a radiation field
 * and J_eff (0 before any).                                                   */
int mgr_features(mgr_ctx *ctx, float *X);
/* Features -> network -> populations (pop_tot 10^y_i / sum 10^y) in place of
 * the current populations; needs 7 levels.  A cell whose prediction is not
 * finite and positive, or gives a line zero or negative opacity
 * (n_j B_ji <= n_i B_ij: an inversion, which a network far outside its
 * training data can predict and on which the iteration diverges), keeps the
 * populations it held before the call (the LTE start, after mgr_first_pass
 * from LTE).                                                                  */
int mgr_warm_start(mgr_ctx *ctx, mgr_mlp *mlp);
/* Cells that kept their previous populations in the last mgr_warm_start.     */
int mgr_warm_start_fallbacks(const mgr_ctx *ctx, int64_t *cells);
/* Whether the network reads the Jlin inputs (4..7; the default).  A network
 * trained with those inputs at 0 (use = 0) needs no first pass: mgr_warm_start
 * then sets them to 0 and needs no radiation field before it.                 */
int mgr_mlp_set_jlin_inputs(mgr_mlp *mlp, int use);
/* The warm start's first pass from the current populations: a radiation
 * field over every pair_stride-th ray pair of each origin (ray weights scaled
 * by pair_stride, an approximation of J at 1/pair_stride of the cost; 1 is
 * mgr_compute_radiation_field) and J_eff, after which mgr_features and
 * mgr_warm_start may be called.  The network must be trained on features of
 * the same stride.                                                            */
int mgr_first_pass(mgr_ctx *ctx, int precision, int pair_stride);

#ifdef __cplusplus
}
#endif

#endif /* MAGRITTE_HIP_H */
