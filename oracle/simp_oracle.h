/*
 * simp_oracle.h -- CPU FP64 restatement of SIMP's SPH viscous-plastic step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path
 * in simp_amd/ is compared against.  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py may load it; the product never does.
 *
 * Layout follows the reference's Fortran module state: x(dim, n) is
 * interleaved (x1, y1, x2, y2, ...); pair and boundary-pair indices are
 * 1-based, exactly as NNPS writes them.
 */
#ifndef SIMP_ORACLE_H
#define SIMP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

/* Forcing and switches of one simulation (the runtime form of the
 * compile-time PARAMETERs the reference uses on this path). */
typedef struct orc_params {
    int    tse;          /* 0: keep dt (time_step.f90:21-22); 1: Monaghan dt (:39-125);
                            2: VP dt (:132-157) */
    int    wind_mode;    /* 0: uniform wind; 1: gridded wind, bilinear sampling */
    double wind[2];      /* uniform wind (m/s); the reference hard-codes (0, -5.3f) */
    double current[2];   /* uniform ocean current (m/s); reference: (0, 0) */
    double max_hsml;     /* smoothing-length cap (initialize.f90:191) */
    int    den_alg;      /* 3: diagnostic rho = h rho_ice (density.f90:172-194);
                            1: summation (density.f90:41-100);
                            2: continuity (density.f90:107-165) */
    int    grid_nx, grid_ny;
    double grid_x0, grid_y0, grid_dx, grid_dy;
    const double *grid_u;  /* grid_ny * grid_nx, row-major (y outer) */
    const double *grid_v;
    int    pa_sph;       /* 1 (or 0): stress_1 (internal_stress.f90:105-136);
                            2: stress_2 (:189-216) */
    int    eos_alg;      /* 3 (or 0): kreyscher_eos (eos.f90:141-147);
                            1: hibler_eos (:59); 2: gutfraind_eos (:94-102) */
    int    rheology;     /* 1 (or 0): elliptical yield curve
                            (bulk_viscosity.f90:35-79, shear_viscosity.f90:36-60);
                            2: Mohr-Coulomb (bulk_viscosity.f90:86-145,
                            shear_viscosity.f90:67-124) */
    int    skf;          /* 6 (or 0): Wendland C6 (kernel.f90:471-520,
                            boundary_kernel.f90:438-504); 4: Wendland C2
                            (kernel.f90:359-408, boundary_kernel.f90:290-357);
                            5: Wendland C4 (kernel.f90:415-464,
                            boundary_kernel.f90:364-431); 1: quadratic
                            (kernel.f90:241-290, boundary_kernel.f90:54-120);
                            2: Gaussian (kernel.f90:108-157,
                            boundary_kernel.f90:127-192); 3: quintic spline
                            (kernel.f90:165-233, boundary_kernel.f90:199-283);
                            7, 8: the Marquis kernels (kernel.f90:527-633)
                            with b_Gaussian (boundary_kernel.f90:38-43) */
    int    int_algorithm; /* 2 (or 0): predictor_corrector
                            (time_integration.f90:155-282); 1: euler_scheme
                            (:86-148); 3: leapfrog_scheme (:289-391) */
    int    sle;          /* smoothing_length (smoothing_length.f90:21-30):
                            2: Marquis_hsml (:118-145); 1: Benz_hsml
                            (:40-111); 0: unchanged.  No zero default. */
    int    xsph;         /* xsph_regularization (globals.f90:187):
                            XSPH (regularization.f90:45-95) */
    int    art_visc;     /* artificial_visc (globals.f90:186) */
    int    art_visc_alg; /* 1 (or 0): art_visc_HK (regularization.f90:102-213);
                            2: art_visc_Mon (:220-308).  Either runs whenever
                            regularization does (:29-35) */
    int    bfa;          /* 1 (or 0): bf_Monaghan (boundary_force.f90:38-102);
                            2: bf_Wang (:109-176); 3: bf_Marquis (:183-223) */
    const double *b_normal;  /* (2, nb) boundary normals, bf_Wang only */
    double step_dt;      /* the module dt single_step sees (bf_Wang reads it);
                            orc_run_pc sets it to each step's dt */
    /* the logical scenario switches of globals.f90:178-187 (no zero
     * default: 1 is the reference's value except for boundary_friction) */
    int    boundary_friction;  /* tangential boundary forces */
    int    boundary_pressure;  /* normal boundary forces; with neither, no
                                  boundary pairs (nnps.f90:440) */
    int    pressure_force;     /* 0: eos is not called, p stays the ice
                                  strength P (internal_stress.f90:35-39) */
    int    wind_force, water_force;
    int    viscosity;          /* 0: zeta, eta stay zero and p is set by eos
                                  only (internal_stress.f90:20-39); with
                                  pressure_force 0 also the strain rate is
                                  never set (single_step.f90:80) */
} orc_params;

orc_ctx *orc_create(void);
void     orc_destroy(orc_ctx *c);

/* bucket_search (nnps.f90:155-517): pair list in the reference's order.
 * Arrays are owned by the context and valid until the next call. */
int orc_nnps(orc_ctx *c, int n, const double *x, const double *hsml,
             int nb, const double *xb);
/* direct_find (nnps.f90:41-148): O(N^2) brute force, same arrays. */
int orc_direct_find(orc_ctx *c, int n, const double *x, const double *hsml,
                    int nb, const double *xb);
int orc_nip(const orc_ctx *c);
int orc_bnip(const orc_ctx *c);
/* Copy the last pair lists out (1-based indices). */
void orc_get_pairs(const orc_ctx *c, int *p_i, int *p_j, double *r_ij,
                   double *hsml_ij);
void orc_get_bpairs(const orc_ctx *c, int *bp_i, int *bp_j, double *br_ij);
/* Reference NNPS bin statistics of the last orc_nnps call:
 * out[0] = bin_size, out[1..4] = bucket_coor (xlo, xhi, ylo, yhi). */
void orc_get_bins(const orc_ctx *c, double *out);
/* Kernel values w(1:nip) and kernel_zero(1:n) of the last orc_single_step. */
int orc_get_kernel(const orc_ctx *c, int n, double *w, double *kernel_zero);

/* single_step (single_step.f90:5-108) on interleaved host arrays.
 * eps3 and sig3 are (3, n): (e11, e12, e22) and (s11, s12, s22). */
int orc_single_step(orc_ctx *c, const orc_params *prm, int n,
                    const double *x, const double *vx, const double *mass,
                    const double *rho, const double *hsml, const double *h,
                    const double *A, int nb, const double *xb,
                    const double *massb, double current_time,
                    double *dxdt, double *dvxdt, double *dhdt, double *dAdt,
                    double *eps3, double *zeta, double *eta, double *p,
                    double *sig3);

/* nsteps of time_integration's loop (time_integration.f90:38-73) with the
 * scheme prm->int_algorithm selects (predictor_corrector by default).  State arrays are updated in place;
 * dt, current_time and itimestep advance as in the reference.  The
 * diagnostic outputs (eps3, zeta, eta, p, sig3) hold the last single_step's
 * values and may be NULL.  With sle = 0 or 1 the masses evolve
 * (mass_evolution.f90:40-71) and mass is updated in place as well. */
int orc_run_pc(orc_ctx *c, const orc_params *prm, int nsteps, int n,
               double *x, double *vx, double *mass, double *rho,
               double *hsml, double *h, double *A, int nb, const double *xb,
               const double *massb, double *dt, double *current_time,
               int *itimestep, double *eps3, double *zeta, double *eta,
               double *p, double *sig3);

/* sum_density (density.f90:41-100) with the kernel values and kernel_zero
 * of the last orc_single_step on this context; rho is updated in place. */
int orc_sum_density(orc_ctx *c, int n, const double *mass, double *rho);

/* Wind at n points (x interleaved, uv interleaved out): uniform, or the
 * bilinear sample of the grid clamped at its edge. */
void orc_wind_at(const orc_params *prm, int n, const double *x, double *uv);

/* VP_time_step (time_step.f90:155). */
/* con_density (density.f90:107-165): drhodt from the last single_step's
 * pairs and gradients with the velocities vx (2, n). */
int orc_con_density(orc_ctx *c, int n, const double *vx, const double *mass,
                    double *drhodt);
/* the pair gradients dwdx (2, nip) of the last single_step */
void orc_get_dwdx(const orc_ctx *c, double *dwdx);
/* c_yang sound speed (sound_speed.f90:36-60) of a density */
double orc_sound_speed(double rho);
double orc_vp_dt(int n, const double *hsml);
/* Monaghan_time_step without artificial viscosity (time_step.f90:39-125)
 * from hsml, rho and the last evaluation's dvxdt (2, n). */
double orc_monaghan_dt(int n, const double *hsml, const double *rho, const double *dvxdt);

#ifdef __cplusplus
}
#endif
#endif
