/*
 * simp.h -- C ABI of the MI355X (gfx950) SPH viscous-plastic step.
 *
 * The boundary is SIMP's `single_step` (integration/single_step.f90:5,
 * explicit interface modules/interfaces.f90:243-251): a subroutine without
 * dummy arguments that reads and writes module state.  Here that state is
 * passed explicitly as plain host pointers in the reference's Fortran layout
 * (x(dim, n) interleaved, 1-based pair indices), so a Fortran caller binds
 * each entry point with iso_c_binding and passes its module arrays as
 * assumed-size actual arguments.
 *
 * Every function returns a status: SIMP_OK or one of SIMP_E_*.  The reference
 * reports the same conditions with WRITE + STOP (neighbor/nnps.f90:415-423,
 * 660-668); the Fortran shim turns a nonzero status back into STOP.
 *
 * All functions are re-entrant across handles.  One handle must not be used
 * from two host threads at once.
 */
#ifndef SIMP_H
#define SIMP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SIMP_OK             0
#define SIMP_E_ARG          1  /* invalid argument or null pointer             */
#define SIMP_E_HIP          2  /* HIP runtime failure                           */
#define SIMP_E_CAPACITY     3  /* a capacity was exceeded (cells, pairs, >200   */
                               /* interactions per particle as nnps.f90:660)    */
#define SIMP_E_NAN          4  /* non-finite position or smoothing length      */
#define SIMP_E_UNSUPPORTED  5  /* a switch value this build does not implement */
#define SIMP_E_NOMEM        6  /* device allocation failed                     */
#define SIMP_E_STATE        7  /* call out of order (e.g. step before upload)  */

#define SIMP_FP64   0          /* all arithmetic in FP64                        */
#define SIMP_MIXED  1          /* pair sums in FP32 with compensated sums       */

#define SIMP_WIND_UNIFORM 0
#define SIMP_WIND_GRID    1

#define SIMP_CONFIG_VERSION 3

/* simp_instance_desc.sle value for the reference's sle = 0 (smoothing
 * lengths kept); 0 there means the default, Marquis_hsml (sle = 2). */
#define SIMP_SLE_KEEP (-1)

/* Runtime form of the compile-time switches of modules/globals.f90:58-187
 * that reach the hot path.  The values each field lists are implemented;
 * any other value is rejected with SIMP_E_UNSUPPORTED. */
typedef struct simp_config {
    int32_t version;        /* SIMP_CONFIG_VERSION                              */
    int32_t precision;      /* SIMP_FP64 or SIMP_MIXED                          */
    int32_t int_algorithm;  /* 2: predictor-corrector (globals.f90:64),         */
                            /* time_integration.f90:155-282; 1: Euler (:86-148) */
                            /* 3: leap-frog (:289-391); batches run 2           */
    int32_t nnps_algorithm; /* 2: bucket search (globals.f90:73)                */
    int32_t sle;            /* 2: hsml = 3 sqrt(m/rho) (globals.f90:81,        */
                            /* smoothing_length.f90:118-145); 1: Benz's rate  */
                            /* (:40-111); 0: kept; with 0 and 1 the masses    */
                            /* evolve (mass_evolution.f90:40-71)              */
    int32_t tse;            /* 0: keep dt, 2: VP dt (globals.f90:89); 1: the    */
                            /* Monaghan dt (time_step.f90:39-125), with the    */
                            /* viscous CFL terms when artificial_visc is on    */
    int32_t rheology;       /* 1: elliptical yield curve (globals.f90:96);      */
                            /* 2: Mohr-Coulomb (bulk_viscosity.f90:86-145,      */
                            /* shear_viscosity.f90:67-124)                      */
    int32_t bfa;            /* 1: Monaghan boundary force (globals.f90:119);    */
                            /* 2: Wang (boundary_force.f90:109-176; reads the  */
                            /* boundary normals and dt, simp_set_module_state);*/
                            /* 3: Marquis (:183-223, no force while the        */
                            /* boundary pressure is on)                        */
    int32_t eos_alg;        /* 3: Kreyscher replacement pressure (:127),        */
                            /* eos.f90:141-147; 1: Hibler P (eos.f90:59);     */
                            /* 2: Gutfraind (eos.f90:94-102)                  */
    int32_t pa_sph;         /* 1: sigma_i/rho_i^2 + sigma_j/rho_j^2 (:134),     */
                            /* internal_stress.f90:105-136 (stress_1);         */
                            /* 2: (sigma_i + sigma_j)/(rho_i rho_j), :189-216  */
    int32_t skf;            /* 6: Wendland C6 (globals.f90:147); 4: Wendland    */
                            /* C2, 5: Wendland C4 (kernel.f90:359-464,          */
                            /* boundary_kernel.f90:290-431); 1: quadratic,      */
                            /* 2: Gaussian, 3: quintic spline, 7, 8: Marquis    */
                            /* (kernel.f90:108-290, 527-633)                    */
    int32_t den_alg;        /* 3: diagnostic density rho = h rho_ice (:155);    */
                            /* 1: summation density (density.f90:41-100);       */
                            /* 2: continuity density (density.f90:107-165)      */
    int32_t wind_mode;      /* SIMP_WIND_UNIFORM or SIMP_WIND_GRID              */
    int32_t max_particles;  /* capacity of particles (0: size at upload)        */
    int32_t max_boundary;   /* capacity of boundary particles (0: at upload)    */
    int32_t reserved;
    double  wind[2];        /* uniform wind (m/s); reference (0, -5.3f),        */
                            /* ext_force/wind_forcing.f90:36-37                 */
    double  current[2];     /* uniform ocean current (m/s); reference (0, 0),   */
                            /* ext_force/water_drag.f90:37-38                   */
    double  max_hsml;       /* smoothing-length cap (in_out/initialize.f90:191) */
    int32_t artificial_visc;      /* 0 / 1 (globals.f90:186)                   */
    int32_t art_visc_alg;         /* 1: Hernquist-Katz (regularization.f90:     */
                                  /* 102-213); 2: Monaghan (:220-308)           */
    int32_t xsph_regularization;  /* 0 / 1: XSPH (regularization.f90:45-95);    */
                                  /* regularization then also runs the          */
                                  /* art_visc_alg viscosity (:29-35)            */
    /* the other logical scenario switches (globals.f90:178-187), 0 / 1;
     * simp_config_default gives the reference's values (friction 0, the
     * others 1) */
    int32_t boundary_friction;    /* tangential boundary forces (boundary_    */
                                  /* force.f90:87-100, 157-166, 217-221)       */
    int32_t boundary_pressure;    /* normal boundary forces; without either    */
                                  /* boundary force no boundary pairs exist    */
                                  /* (nnps.f90:440) and the boundary is unused */
    int32_t pressure_force;       /* 0: eos is not called, p stays the ice     */
                                  /* strength P (internal_stress.f90:35-39)    */
    int32_t wind_force;           /* wind_forcing.f90:35                       */
    int32_t water_force;          /* water_drag.f90:35                         */
    int32_t viscosity;            /* 0 with pressure_force 0: the strain rate  */
                                  /* is never set and stays zero (single_step. */
                                  /* f90:80); with pressure_force 1 no effect  */
} simp_config;

typedef struct simp_handle simp_handle;

/* Fill *cfg with the reference's active switch values (globals.f90). */
void simp_config_default(simp_config *cfg);

/* Create one simulation context on the current HIP device. */
int simp_create(const simp_config *cfg, simp_handle **out);
int simp_destroy(simp_handle *h);

/* Static wind field on a regular grid (row-major, ny rows of nx values),
 * sampled bilinearly at particle positions when wind_mode is
 * SIMP_WIND_GRID.  The reference has a uniform field only
 * (wind_forcing.f90:36-37); the grid carries the synthetic forcing of the
 * larger workloads. */
int simp_set_wind_grid(simp_handle *h, int nx, int ny, double x0, double y0,
                       double dx, double dy, const double *u, const double *v);

/* ---- the drop-in for single_step (integration/single_step.f90:5-108) ----
 * One right-hand-side evaluation on host arrays: copies the state in, runs
 * NNPS, the pair sums, the stress, boundary and forcing terms on the GPU and
 * copies the results out.  x, vx, dxdt, dvxdt are (2, n) interleaved; eps3 =
 * (epsilon_11, epsilon_12, epsilon_22) and sig3 = (sigma_11, sigma_12,
 * sigma_22) are (3, n) interleaved.  current_time gates wind and water drag
 * (wind_forcing.f90:35, water_drag.f90:36); itimestep selects the
 * interaction-count check of nnps.f90:512-515.  Any output pointer may be
 * NULL.  nip and bnip receive the pair counts NNPS would hold. */
int simp_single_step(simp_handle *h, int n, const double *x, const double *vx,
                     const double *mass, const double *rho, const double *hsml,
                     const double *hth, const double *A, int nb,
                     const double *xb, const double *massb,
                     double current_time, int itimestep,
                     double *dxdt, double *dvxdt, double *dhdt, double *dAdt,
                     double *eps3, double *zeta, double *eta, double *p,
                     double *sig3, int *nip, int *bnip);

/* ---- simp_single_step without allocation ----
 * No counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and status codes as
 * simp_single_step, but the call allocates nothing, on the device or on
 * the host: the device state lives in one array per GPU that the library
 * holds from its first simp_create on that GPU until the process ends
 * (SIMP_NOALLOC_SCRATCH_BYTES: one instance of up to about 250,000
 * particles at the reference's spacing; the hsml_ij store holds 262,144
 * boundary pairs).  An input that needs more returns SIMP_E_CAPACITY
 * before anything is touched.
 * Every handle on one GPU uses the same array, which keeps the last call's
 * contents: the hsml_ij store and kernel_zero persist from call to call as
 * the reference's module arrays do (a handle's first call starts them from
 * zero, as a new handle's simp_single_step does), shared by all the
 * handles on that GPU like a module array shared by all callers in one
 * process.  Calls that overlap on one GPU (threads, several handles) race.
 * simp_kernel_values and simp_interaction_stats read the last call back
 * while no other handle's call has followed it (SIMP_E_STATE otherwise).
 * A handle that holds a layout from the allocating entry points returns
 * SIMP_E_STATE here. */
#define SIMP_NOALLOC_SCRATCH_BYTES (256u << 20)
int simp_single_step_noalloc(simp_handle *h, int n, const double *x,
                             const double *vx, const double *mass,
                             const double *rho, const double *hsml,
                             const double *hth, const double *A, int nb,
                             const double *xb, const double *massb,
                             double current_time, int itimestep,
                             double *dxdt, double *dvxdt, double *dhdt,
                             double *dAdt, double *eps3, double *zeta,
                             double *eta, double *p, double *sig3, int *nip,
                             int *bnip);

/* ---- module state single_step reads besides its arguments ----
 * The Wang boundary force (bfa = 2, boundary_force.f90:145-174) reads the
 * time step dt and the boundary normals b_normal (data_variables.f90:41-46,
 * set by boundary_placing_1, in_out/boundaries.f90:232-287): (2, nb)
 * interleaved unit vectors pointing into the ice.  simp_single_step and
 * simp_upload take them from the last call of this function (dt: only
 * simp_single_step; simp_upload has its own).  b_normal may be NULL (zeros);
 * nb must then be 0 or match the next call's boundary count.  The sound
 * speed the artificial viscosity and the Monaghan time step read is c_yang
 * of the given densities, as sound_speed leaves it before every
 * single_step call of the reference (initialize.f90:199-200,
 * time_integration.f90:140, 230, 274, 355, 383). */
int simp_set_module_state(simp_handle *h, double dt, int nb, const double *b_normal);

/* ---- kernel values of the last single_step, for a host-side density ----
 * Replaces the module arrays single_step leaves for sum_density
 * (phy_quantities/density.f90:41-100; written at sph/kernel.f90:503,511):
 * the particle pairs of the last simp_single_step in the reference's order
 * (p_i, p_j 1-based), their Wendland C6 values w (kernel.f90:506), and
 * kernel_zero(1:n), the factor of the last pair each particle is the first
 * member of; it keeps its value across calls for particles that are first
 * in no pair, as the module array does, and is tracked only when the
 * handle's den_alg is 1 (SIMP_E_STATE otherwise).  Any array may be NULL;
 * nip is always written; SIMP_E_CAPACITY when it exceeds cap. */
int simp_kernel_values(simp_handle *h, int cap, int *nip, int *p_i, int *p_j,
                       double *w, double *kernel_zero);

/* ---- kernel gradients of the last single_step, for a host-side density --
 * Replaces the module arrays single_step leaves for con_density
 * (phy_quantities/density.f90:107-165): the particle pairs of the last
 * simp_single_step in the reference's order (p_i, p_j 1-based) and the
 * kernel gradient dwdx of each (written by the kernel routines of
 * sph/kernel.f90, e.g. :507), in the pair's orientation dx = x(p_i) -
 * x(p_j), (2, nip) interleaved.  Any array may be NULL; nip is always
 * written; SIMP_E_CAPACITY when it exceeds cap (pairs). */
int simp_kernel_gradients(simp_handle *h, int cap, int *nip, int *p_i, int *p_j,
                          double *dwdx);

/* ---- how often the neighbour search runs ----
 * The reference's bucket_search (nnps.f90:155-517) sorts the particles into
 * cells and tests the 3x3 block of every particle at every evaluation.  This
 * library rebuilds its cell lists only in evaluations where a particle
 * changed cell, and its candidate lists (the 3x3 block within mhsml plus a
 * skin) only when the particles may have moved by more than the skin; the
 * pair sets and their order are the same either way.
 * simp_search_stats: the evaluations of the handle's instance since its
 * upload, and of them those that rebuilt the cell lists and those that
 * rebuilt the candidate lists.  simp_set_search(h, 1) rebuilds both at
 * every evaluation (same results; a measurement of the full search),
 * simp_set_search(h, 0) returns to the default. */
int simp_set_search(simp_handle *h, int every_evaluation);
int simp_search_stats(simp_handle *h, long long *evaluations,
                      long long *cell_rebuilds, long long *candidate_rebuilds);

/* ---- tuning and the autotuner (no counterpart in the reference) ----
 * Parameters of the step kernels that change no result, only the speed: the
 * neighbours whose loads pass A's exact pair test (passA_batch, 1-4) and
 * pass B's sums (passB_batch, 2-4; 0: 3 in mixed precision, 4 in FP64)
 * issue together, and the skin of the candidate lists as a fraction of
 * MINVAL(hsml) (0 < skin <= 0.2).  The cell size stays the reference's
 * 2 MINVAL(hsml) (nnps.f90:211): the cells fix the pair order.
 * simp_tune / simp_batch_tune time each candidate setting over
 * steps_per_trial steps of the resident content, one parameter at a time
 * (pass A's batch, pass B's, the skin), and keep the fastest; the trial
 * steps advance the simulation (about 11 trials).  us_start and us_best
 * receive the microseconds per step of the setting it started from and of
 * the one it keeps; either may be NULL. */
typedef struct simp_tuning {
    int32_t passA_batch;
    int32_t passB_batch;
    int32_t reserved;
    double  skin;
} simp_tuning;
int simp_set_tuning(simp_handle *h, const simp_tuning *t);
int simp_get_tuning(simp_handle *h, simp_tuning *t);
int simp_tune(simp_handle *h, int steps_per_trial, simp_tuning *chosen,
              double *us_start, double *us_best);

/* ---- interaction_stats (neighbor/nnps.f90:613-658) ----
 * The neighbour statistics the reference prints from NNPS at every step that
 * is a multiple of print_step (nnps.f90:512-515): the largest and smallest
 * number of interactions of one particle (particle pairs plus boundary pairs
 * it is in), their average over the particles (REAL(4) division as
 * nnps.f90:651), the number of particles with none, and the number of
 * particle pairs, of the last evaluation that collected them -- every
 * evaluation of such a step and every simp_single_step call; itimestep
 * receives that evaluation's step.  Counts above 4095 for one particle are
 * reported as 4095.  SIMP_E_STATE before any collecting evaluation.  The
 * >200-interactions STOP of nnps.f90:660-668 is the SIMP_E_CAPACITY status
 * of the stepping calls. */
int simp_interaction_stats(simp_handle *h, int *maxip, int *minip, double *avgip,
                           int *noip, long long *nip, int *itimestep);

/* ---- NNPS (neighbor/nnps.f90:21-30 -> bucket_search :155-517) ----
 * The pair lists in the reference's exact order (single-thread order):
 * p_i, p_j 1-based, r_ij, hsml_ij; boundary pairs bp_i (boundary particle),
 * bp_j (particle), br_ij.  cap / bcap are the array capacities; the counts
 * are always written, and SIMP_E_CAPACITY is returned if they exceed them. */
int simp_nnps(simp_handle *h, int n, const double *x, const double *hsml,
              int nb, const double *xb, int cap, int *nip, int *p_i, int *p_j,
              double *r_ij, double *hsml_ij, int bcap, int *bnip, int *bp_i,
              int *bp_j, double *br_ij);

/* ---- device-resident time integration ----
 * Upload a state once, then advance it by whole predictor-corrector steps
 * (time_integration.f90:38-73 with predictor_corrector :155-282) without
 * host round trips, and download it when needed. */
int simp_upload(simp_handle *h, int n, const double *x, const double *vx,
                const double *mass, const double *rho, const double *hsml,
                const double *hth, const double *A, int nb, const double *xb,
                const double *massb, double dt, double current_time,
                int itimestep);
/* Advance nsteps steps (asynchronous on the handle's stream unless sync). */
int simp_step_pc(simp_handle *h, int nsteps, int sync);
int simp_download(simp_handle *h, double *x, double *vx, double *rho,
                  double *hsml, double *hth, double *A, double *dt,
                  double *current_time, int *itimestep);
/* Diagnostic fields of the last corrector evaluation (written by
 * simp_step_pc only when enabled with simp_set_diagnostics). */
int simp_set_diagnostics(simp_handle *h, int enable);
int simp_download_diagnostics(simp_handle *h, double *eps3, double *zeta,
                              double *eta, double *p, double *sig3);
/* Wait for the handle's stream; returns the first device-side error. */
int simp_sync(simp_handle *h);
/* The HIP stream the handle launches on (hipStream_t). */
void *simp_stream(simp_handle *h);

/* Per-kernel timing of nsteps steps: HIP events bracket every launch on the
 * handle's stream.  names: '\n'-separated kernel names written to buf;
 * ms: average duration per launch of each name; counts: launches per name.
 * Returns the number of names written (at most maxk). */
int simp_profile_steps(simp_handle *h, int nsteps, int maxk, char *buf,
                       int buflen, double *ms, int *counts);

const char *simp_strerror(int status);

/* Exact snapshot of a device-resident run: everything simp_upload takes
 * (state in reference order, masses, boundary particles, dt, current_time,
 * itimestep) plus the persistent arrays the next steps read -- the
 * hsml_ij store that the boundary kernel reads stale and kernel_zero -- so
 * that a handle restored with simp_snapshot_load (same configuration)
 * continues bit for bit as the original would have.  The reference has no
 * counterpart: its only resume path is the CSV spin-up reader (8 digits,
 * in_out/initialize.f90:234-330), which cannot resume exactly.
 * simp_snapshot_size returns the blob's bytes (or -status). */
long long simp_snapshot_size(simp_handle *h);
int simp_snapshot_save(simp_handle *h, void *buf, long long cap);
int simp_snapshot_load(simp_handle *h, const void *buf, long long len);
/* Header of a blob (host only, no device): its particle and boundary
 * counts, step and time, after checking that the blob of len bytes holds
 * everything the header announces -- what simp_snapshot_load checks first;
 * a caller can size max_particles / max_boundary from it.  SIMP_E_ARG for a
 * blob that is not a snapshot, is truncated, or announces impossible
 * counts.  Any output pointer may be NULL. */
int simp_snapshot_info(const void *buf, long long len, int *n, int *nb, int *itimestep,
                       double *current_time);

/* ---- snapshots in the reference's CSV format (host only, no device) ----
 * The files of a run folder: particle_momentum.csv, particle_state.csv and
 * particle_other.csv.  simp_output_create empties (or creates) them, as
 * the reference's initialize does (in_out/initialize.f90:83-94).
 * simp_output_append writes one snapshot exactly as the reference's output
 * routine does (in_out/output.f90:47-78, formats :61-63): per particle i
 * the position in km and the velocity in cm/s; mass, p, hsml, A, h,
 * epsilon_11/12/22, sigma_11/12/22; then n, nb and current_time in hours.
 * Arrays are in the Fortran layout (x, vx, eps3, sig3: component-fastest).
 * simp_output_read_last reads the last snapshot of a folder back: the
 * reference's spin-up reader (in_out/initialize.f90:234-330) with the
 * column layout the writer produces (the reference's reader expects 14
 * state columns and reads the first snapshot); positions and velocities
 * come back in m and m/s, current_time in s.  Returns SIMP_E_CAPACITY (n
 * set) when the snapshot holds more than cap particles. */
int simp_output_create(const char *folder);
int simp_output_append(const char *folder, int n, const double *x, const double *vx,
                       const double *mass, const double *p, const double *hsml,
                       const double *A, const double *hth, const double *eps3,
                       const double *sig3, int nb, double current_time);
int simp_output_read_last(const char *folder, int cap, int *n, int *nb,
                          double *current_time, double *x, double *vx, double *mass,
                          double *p, double *hsml, double *A, double *hth,
                          double *eps3, double *sig3);

/* ---- ensembles: many independent simulations in one packed launch ----
 * No reference counterpart (the reference runs one simulation per process,
 * modules/globals.f90:21-25).  Every step of a batch advances all its active
 * instances with one launch per kernel; instances of different sizes, domains
 * and forcing are packed with per-instance offsets.  An instance freezes when
 * its step budget is used up or when it fails (its status says why); the
 * others continue.  Adding or removing instances rebuilds the packed layout
 * on the device (live state is copied device to device). */
typedef struct simp_batch simp_batch;

typedef struct simp_instance_desc {
    int32_t n, nb;                       /* particles, boundary particles      */
    const double *x, *vx, *mass, *rho, *hsml, *hth, *A;   /* as simp_upload   */
    const double *xb, *massb;
    double  dt, current_time;
    int32_t itimestep;
    int32_t nsteps;                      /* steps to run, then freeze; 0: none */
    int32_t tse;                         /* 0 (keep dt), 1 (Monaghan), 2 (VP)   */
    int32_t wind_mode;                   /* SIMP_WIND_UNIFORM / SIMP_WIND_GRID */
    double  wind[2], current[2], max_hsml;
    int32_t grid_nx, grid_ny;            /* wind grid (wind_mode = GRID)       */
    double  grid_x0, grid_y0, grid_dx, grid_dy;
    const double *grid_u, *grid_v;
    int32_t den_alg;                     /* 0 or 3: diagnostic; 1: summation;  */
                                         /* 2: continuity                      */
    int32_t pa_sph;                      /* 0 or 1: stress_1; 2: stress_2      */
    int32_t eos_alg;                     /* 0 or 3: Kreyscher; 1: Hibler; 2:   */
                                         /* Gutfraind                          */
    int32_t rheology;                    /* 0 or 1: elliptical; 2: Mohr-Coulomb*/
    int32_t skf;                         /* 0 or 6: Wendland C6; 1-5, 7, 8 as  */
                                         /* simp_config.skf                    */
    int32_t reserved;
    int32_t sle;                         /* 0 or 2: Marquis hsml; 1: Benz;     */
                                         /* SIMP_SLE_KEEP: kept (sle = 0)      */
    int32_t xsph;                        /* 1: XSPH regularization             */
    int32_t art_visc;                    /* 1: artificial viscosity            */
    int32_t art_visc_alg;                /* 0 or 1: Hernquist-Katz; 2: Monaghan*/
    int32_t bfa;                         /* 0 or 1: Monaghan; 2: Wang; 3:      */
                                         /* Marquis boundary force             */
    int32_t boundary_friction;           /* 1: boundary friction               */
    const double *b_normal;              /* (2, nb) boundary normals (Wang),   */
                                         /* or NULL (zeros)                    */
    int32_t forces_off;                  /* SIMP_OFF_* bits: the reference's   */
                                         /* logical switches set to .FALSE.    */
    int32_t reserved2;
} simp_instance_desc;

/* simp_instance_desc.forces_off bits */
#define SIMP_OFF_BOUNDARY_PRESSURE 1
#define SIMP_OFF_PRESSURE_FORCE    2
#define SIMP_OFF_WIND              4
#define SIMP_OFF_WATER             8
#define SIMP_OFF_VISCOSITY        16

int simp_batch_create(int precision, simp_batch **out);
int simp_batch_destroy(simp_batch *b);
/* Add count instances; ids[k] receives the id of desc[k]. */
int simp_batch_add(simp_batch *b, int count, const simp_instance_desc *desc, int *ids);
int simp_batch_remove(simp_batch *b, int count, const int *ids);
/* Remove nremove instances and add nadd new ones (ids receives their ids)
 * with one rebuild of the packed layout: what a request loop does at a
 * step boundary where some simulations finish and others are admitted. */
int simp_batch_update(simp_batch *b, int nremove, const int *remove_ids, int nadd,
                      const simp_instance_desc *desc, int *ids);
/* Advance every active instance by nsteps predictor-corrector steps. */
int simp_batch_step(simp_batch *b, int nsteps, int sync);
int simp_batch_sync(simp_batch *b);
/* itimestep, steps still to run (< 0: no limit), active flag, and the
 * instance's status (SIMP_OK, SIMP_E_NAN, SIMP_E_CAPACITY). */
int simp_batch_status(simp_batch *b, int id, int *itimestep, int *steps_left, int *active,
                      int *status);
int simp_batch_download(simp_batch *b, int id, double *x, double *vx, double *rho,
                        double *hsml, double *hth, double *A, double *dt,
                        double *current_time, int *itimestep);
int simp_batch_count(simp_batch *b, int *instances, long long *particles);
/* Status of every instance in one call (one device-to-host copy): writes up
 * to cap entries and returns the number of instances (or -status). */
int simp_batch_list(simp_batch *b, int cap, int *ids, int *itimestep, int *steps_left,
                    int *active, int *status);
/* simp_set_tuning / simp_tune for a batch (all its instances at once). */
int simp_batch_set_tuning(simp_batch *b, const simp_tuning *t);
int simp_batch_tune(simp_batch *b, int steps_per_trial, simp_tuning *chosen,
                    double *us_start, double *us_best);
/* simp_interaction_stats of one instance of the batch. */
int simp_batch_interaction_stats(simp_batch *b, int id, int *maxip, int *minip,
                                 double *avgip, int *noip, long long *nip,
                                 int *itimestep);
/* Device bytes an instance would occupy (for admission against a budget). */
long long simp_instance_footprint(const simp_instance_desc *desc);

#ifdef __cplusplus
}
#endif
#endif
