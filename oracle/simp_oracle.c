/*
 * simp_oracle.c -- CPU FP64 restatement of SIMP's single_step hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Nothing in simp_amd/ links or
 * calls this file; tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg are its only users.
 *
 * Every function restates one reference routine, cited file:line relative to
 * the SIMP source tree (fortran/src/...).  Arithmetic is written in the
 * reference's evaluation order, built with -ffp-contract=off, so that a
 * single-thread reference build and this file agree to rounding of libm
 * calls (exp, sqrt) and of the compiler's integer-power expansion.
 *
 * Parity pins (all in tests/):
 *   - nnps / direct_find: golden vectors written by the compiled reference
 *     (tests/golden) and the closed-form set predicate on brute force.
 *   - single_step, run_pc: golden vectors of the compiled reference, plus
 *     closed-form invariants (uniform velocity => zero strain rate, rest =>
 *     zero stress, stress-pair antisymmetry => zero net internal force).
 *   - sum_density (den_alg = 1): golden vectors of the reference built with
 *     that PARAMETER, including the stale kernel_zero reads, plus the
 *     uniform-density fixed point.
 *   - con_density (den_alg = 2): golden vectors of the reference built with
 *     that PARAMETER, plus a zero rate for a uniform translation and
 *     -rho div(v) for a uniform expansion on a lattice.
 *   - sle = 0 / 1, XSPH, the artificial viscosities, bf_Wang and bf_Marquis:
 *     golden vectors of the reference built with each PARAMETER, plus
 *     closed forms (XSPH and the viscosities vanish for uniform flow; a
 *     single diverging pair gets no viscosity; Wang's force cancels the
 *     normal velocity in one dt; Benz's rate is -h/2 div v / rho for a
 *     uniform expansion).
 */
#include "simp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- constants: globals.f90:199, data_variables.f90:153-180 ---------- */
/* Several reference constants are default-REAL literals widened to REAL(8);
 * (double)(float) reproduces that widening exactly. */
static const double PI_ = 3.1415926535897932;
static const double P_STAR = 2.75e4;
static const double C_EOS = 20.0;
static const double RHO_ICE = 900.0;
static const double RHO_WATER = 1026.0;
static const double C_WIND = 1.2e-3;
static const double C_WATER = 5.5e-3;
static const double DELTA_MIN = 2.0e-9;
static const double B_LENGTH = 8000.0;
static const double KAPPA_N = 5.4e12;
static const double K_T = 0.0;            /* T_star / P_star */
static const double E_ELL = 2.0;          /* 0.5 (P*+T*) / S*, exact */
static const double HSML_MULT = 3.0;      /* hsml_multiplicity */
#define RHO_AIR ((double)1.3f)            /* rho_air = 1.3 (REAL(4) literal) */
/* Mohr-Coulomb (data_variables.f90:153-180): mu = TAN(phi pi / 180.) with
 * phi = 45, 1 - 2^-53 as the reference build folds it; zeta_max and eta_max
 * are declared there but never given a value, and the compiled reference's
 * module storage holds zero, so the caps are zero times the pressure. */
static const double MU_MC = 0x1.fffffffffffffp-1;
static const double ZETA_MAX = 0.0, ETA_MAX = 0.0;
static const double ZERO_MC = 1.e-18;      /* zero = 1.d-18 */
/* artificial viscosity, XSPH and boundary friction (data_variables.f90:
 * 153-180): alpha_pi = 1., beta_pi = 2. (exact), chi = 0.3 and xi = 0.5225
 * default-REAL literals */
static const double ALPHA_PI = 1.0, BETA_PI = 2.0;
#define CHI_XSPH ((double)0.3f)
#define XI_FRIC ((double)0.5225f)
static const double KAPPA_T = 5.e-8;      /* kappa_t = 5.D-8 */

enum { ORC_OK = 0, ORC_E_ARG = 1, ORC_E_NOMEM = 2 };

struct orc_ctx {
    /* particle-pair list (data_variables.f90:81-90) */
    int nip, cap;
    int *p_i, *p_j;
    double *r_ij, *hsml_ij;   /* hsml_ij keeps stale entries beyond nip */
    double *dwdx;             /* (2, nip) */
    double *w;                /* kernel values w(k) (kernel.f90:511) */
    /* kernel_zero(maxn) (kernel.f90:503): the factor of the last pair each
     * particle was the first member of; persists across calls */
    int nkz;
    double *kz;
    /* boundary-pair list */
    int bnip, bcap;
    int *bp_i, *bp_j;
    double *br_ij;
    double bins_info[5];
    /* scratch */
    int nscr;
    double *div, *grad;
    double *aux;              /* (5, n): c, lv / div_v, xsph_sum (2), spare */
    /* drhodt(maxn) (data_variables.f90): Benz_hsml adds each step's pair
     * sum to it (smoothing_length.f90:76-98) and nothing but con_density
     * resets it, so without continuity density it accumulates over the
     * run; zero at the start like the module array (initialize.f90:44) */
    int ndr;
    double *drp;
};

orc_ctx *orc_create(void) { return (orc_ctx *)calloc(1, sizeof(orc_ctx)); }

void orc_destroy(orc_ctx *c) {
    if (!c) return;
    free(c->p_i); free(c->p_j); free(c->r_ij); free(c->hsml_ij); free(c->dwdx);
    free(c->w); free(c->kz);
    free(c->bp_i); free(c->bp_j); free(c->br_ij);
    free(c->div); free(c->grad); free(c->aux); free(c->drp);
    free(c);
}

int orc_nip(const orc_ctx *c) { return c->nip; }
int orc_bnip(const orc_ctx *c) { return c->bnip; }

void orc_get_pairs(const orc_ctx *c, int *p_i, int *p_j, double *r_ij,
                   double *hsml_ij) {
    size_t k = (size_t)c->nip;
    if (p_i) memcpy(p_i, c->p_i, k * sizeof(int));
    if (p_j) memcpy(p_j, c->p_j, k * sizeof(int));
    if (r_ij) memcpy(r_ij, c->r_ij, k * sizeof(double));
    if (hsml_ij) memcpy(hsml_ij, c->hsml_ij, k * sizeof(double));
}

void orc_get_bpairs(const orc_ctx *c, int *bp_i, int *bp_j, double *br_ij) {
    size_t k = (size_t)c->bnip;
    if (bp_i) memcpy(bp_i, c->bp_i, k * sizeof(int));
    if (bp_j) memcpy(bp_j, c->bp_j, k * sizeof(int));
    if (br_ij) memcpy(br_ij, c->br_ij, k * sizeof(double));
}

void orc_get_bins(const orc_ctx *c, double *out) {
    memcpy(out, c->bins_info, sizeof(c->bins_info));
}

int orc_get_kernel(const orc_ctx *c, int n, double *w, double *kernel_zero) {
    if (n > c->nkz) return ORC_E_ARG;
    if (w && c->nip) memcpy(w, c->w, (size_t)c->nip * sizeof(double));
    if (kernel_zero) memcpy(kernel_zero, c->kz, (size_t)n * sizeof(double));
    return ORC_OK;
}

/* Grow the pair arrays; new hsml_ij entries start at 0 like the zeroed
 * module array (initialize.f90:58), old entries are kept (stale reads). */
static int grow_pairs(orc_ctx *c, int need) {
    if (need <= c->cap) return 0;
    int nc = c->cap ? c->cap : 1024;
    while (nc < need) nc *= 2;
    int *a = (int *)realloc(c->p_i, (size_t)nc * sizeof(int));
    if (!a) return -1;
    c->p_i = a;
    a = (int *)realloc(c->p_j, (size_t)nc * sizeof(int));
    if (!a) return -1;
    c->p_j = a;
    double *d = (double *)realloc(c->r_ij, (size_t)nc * sizeof(double));
    if (!d) return -1;
    c->r_ij = d;
    d = (double *)realloc(c->hsml_ij, (size_t)nc * sizeof(double));
    if (!d) return -1;
    c->hsml_ij = d;
    memset(c->hsml_ij + c->cap, 0, (size_t)(nc - c->cap) * sizeof(double));
    d = (double *)realloc(c->dwdx, (size_t)nc * 2 * sizeof(double));
    if (!d) return -1;
    c->dwdx = d;
    d = (double *)realloc(c->w, (size_t)nc * sizeof(double));
    if (!d) return -1;
    c->w = d;
    c->cap = nc;
    return 0;
}

static int grow_bpairs(orc_ctx *c, int need) {
    if (need <= c->bcap) return 0;
    int nc = c->bcap ? c->bcap : 1024;
    while (nc < need) nc *= 2;
    int *a = (int *)realloc(c->bp_i, (size_t)nc * sizeof(int));
    if (!a) return -1;
    c->bp_i = a;
    a = (int *)realloc(c->bp_j, (size_t)nc * sizeof(int));
    if (!a) return -1;
    c->bp_j = a;
    double *d = (double *)realloc(c->br_ij, (size_t)nc * sizeof(double));
    if (!d) return -1;
    c->br_ij = d;
    c->bcap = nc;
    return 0;
}

static int push_pair(orc_ctx *c, int i, int j, double d, double mh) {
    if (c->nip + 1 > c->cap && grow_pairs(c, c->nip + 1)) return -1;
    c->p_i[c->nip] = i; c->p_j[c->nip] = j;
    c->r_ij[c->nip] = d; c->hsml_ij[c->nip] = mh;
    c->nip++;
    return 0;
}

static int push_bpair(orc_ctx *c, int b, int j, double d) {
    if (c->bnip + 1 > c->bcap && grow_bpairs(c, c->bnip + 1)) return -1;
    c->bp_i[c->bnip] = b; c->bp_j[c->bnip] = j; c->br_ij[c->bnip] = d;
    c->bnip++;
    return 0;
}

/* x**k for a small integer constant k, expanded as the reference build does
 * it: a left-to-right product x*x*...*x (k factors). */
static double ipow(double x, int k) {
    double r = x;
    for (int i = 1; i < k; ++i) r = r * x;
    return r;
}

/* Fortran NINT of a REAL(8): round half away from zero. */
static int nint_(double v) { return (int)lround(v); }

/* SQRT(SUM((a - b)**2)) for dim = 2 (nnps.f90:277 and siblings). */
static double dist2d(const double *a, const double *b) {
    double dx = a[0] - b[0];
    double dy = a[1] - b[1];
    return sqrt(dx * dx + dy * dy);
}

/* ---- bucket_search, nnps.f90:155-517 (single thread order) ----------- */
int orc_nnps(orc_ctx *c, int n, const double *x, const double *hsml, int nb,
             const double *xb) {
    if (n < 1 || nb < 0) return ORC_E_ARG;
    c->nip = 0;
    c->bnip = 0;
    /* bin_size = 2 * MINVAL(hsml) (:211) */
    double hmin = hsml[0];
    for (int i = 1; i < n; ++i) if (hsml[i] < hmin) hmin = hsml[i];
    double bin = 2 * hmin;
    /* bucket_coor over [xb, x] (:212-216) */
    double mnx = x[0], mxx = x[0], mny = x[1], mxy = x[1];
    for (int i = 0; i < n; ++i) {
        if (x[2 * i] < mnx) mnx = x[2 * i];
        if (x[2 * i] > mxx) mxx = x[2 * i];
        if (x[2 * i + 1] < mny) mny = x[2 * i + 1];
        if (x[2 * i + 1] > mxy) mxy = x[2 * i + 1];
    }
    for (int b = 0; b < nb; ++b) {
        if (xb[2 * b] < mnx) mnx = xb[2 * b];
        if (xb[2 * b] > mxx) mxx = xb[2 * b];
        if (xb[2 * b + 1] < mny) mny = xb[2 * b + 1];
        if (xb[2 * b + 1] > mxy) mxy = xb[2 * b + 1];
    }
    int c1 = nint_(mnx / bin), c2 = nint_(mxx / bin);
    int c3 = nint_(mny / bin), c4 = nint_(mxy / bin);
    c->bins_info[0] = bin;
    c->bins_info[1] = c1; c->bins_info[2] = c2;
    c->bins_info[3] = c3; c->bins_info[4] = c4;
    /* number_bin(c1-1:c2+1, c3-1:c4+1) and bins(:, ...) (:217-224);
     * the reference's per-cell capacity max_ind_int is not enforced here. */
    long W = (long)(c2 - c1) + 3, H = (long)(c4 - c3) + 3;
    long ncell = W * H;
    int *cnt = (int *)calloc((size_t)ncell + 1, sizeof(int));
    int *start = (int *)calloc((size_t)ncell + 1, sizeof(int));
    int *list = (int *)malloc((size_t)n * sizeof(int));
    int *fill = (int *)calloc((size_t)ncell + 1, sizeof(int));
    if (!cnt || !start || !list || !fill) {
        free(cnt); free(start); free(list); free(fill);
        return ORC_E_NOMEM;
    }
#define CELL(k, l) (((long)(l) - (c3 - 1)) * W + ((long)(k) - (c1 - 1)))
    for (int i = 0; i < n; ++i) {
        int xb_ = nint_(x[2 * i] / bin), yb_ = nint_(x[2 * i + 1] / bin);
        cnt[CELL(xb_, yb_)]++;
    }
    for (long q = 0; q < ncell; ++q) start[q + 1] = start[q] + cnt[q];
    /* serial fill in ascending particle index (:246-254) */
    for (int i = 0; i < n; ++i) {
        long q = CELL(nint_(x[2 * i] / bin), nint_(x[2 * i + 1] / bin));
        list[start[q] + fill[q]++] = i;
    }
    int rc = ORC_OK;
    /* cell loops (:259-411): l = y outer, k = x inner */
    for (int l = c3; l <= c4 && rc == ORC_OK; ++l) {
        for (int k = c1; k <= c2 && rc == ORC_OK; ++k) {
            long q0 = CELL(k, l);
            int s0 = start[q0], n0 = cnt[q0];
            /* within the cell, i < j in bin order (:265-297) */
            for (int a = 0; a < n0 - 1; ++a) {
                int i = list[s0 + a];
                for (int b = a + 1; b < n0; ++b) {
                    int j = list[s0 + b];
                    double d = dist2d(x + 2 * i, x + 2 * j);
                    double mh = (hsml[i] + hsml[j]) / 2.;
                    if (d <= mh && push_pair(c, i + 1, j + 1, d, mh)) rc = ORC_E_NOMEM;
                }
            }
            /* right, up, up-right, up-left neighbours (:300-407) */
            const long nq[4] = {CELL(k + 1, l), CELL(k, l + 1),
                                CELL(k + 1, l + 1), CELL(k - 1, l + 1)};
            for (int a = 0; a < n0; ++a) {
                int i = list[s0 + a];
                for (int t = 0; t < 4; ++t) {
                    int s1 = start[nq[t]], n1 = cnt[nq[t]];
                    for (int b = 0; b < n1; ++b) {
                        int j = list[s1 + b];
                        double d = dist2d(x + 2 * i, x + 2 * j);
                        double mh = (hsml[i] + hsml[j]) / 2.;
                        if (d <= mh && push_pair(c, i + 1, j + 1, d, mh)) rc = ORC_E_NOMEM;
                    }
                }
            }
        }
    }
    /* boundary pairs (:451-493): 3x3 cells around each boundary particle,
     * y outer, x inner, bin order within a cell */
    for (int b = 0; b < nb && rc == ORC_OK; ++b) {
        int xbin = nint_(xb[2 * b] / bin), ybin = nint_(xb[2 * b + 1] / bin);
        for (int kk = ybin - 1; kk <= ybin + 1; ++kk) {
            for (int jj = xbin - 1; jj <= xbin + 1; ++jj) {
                long q = CELL(jj, kk);
                for (int t = 0; t < cnt[q]; ++t) {
                    int j = list[start[q] + t];
                    double d = dist2d(xb + 2 * b, x + 2 * j);
                    if (d <= B_LENGTH && push_bpair(c, b + 1, j + 1, d)) rc = ORC_E_NOMEM;
                }
            }
        }
    }
#undef CELL
    free(cnt); free(start); free(list); free(fill);
    return rc;
}

/* ---- direct_find, nnps.f90:41-148 ------------------------------------ */
int orc_direct_find(orc_ctx *c, int n, const double *x, const double *hsml,
                    int nb, const double *xb) {
    if (n < 1 || nb < 0) return ORC_E_ARG;
    c->nip = 0;
    c->bnip = 0;
    for (int i = 0; i < n - 1; ++i)
        for (int j = i + 1; j < n; ++j) {
            double d = dist2d(x + 2 * i, x + 2 * j);
            double mh = (hsml[i] + hsml[j]) / 2.;
            if (d <= mh && push_pair(c, i + 1, j + 1, d, mh)) return ORC_E_NOMEM;
        }
    for (int b = 0; b < nb; ++b)
        for (int j = 0; j < n; ++j) {
            double d = dist2d(xb + 2 * b, x + 2 * j);
            if (d <= B_LENGTH && push_bpair(c, b + 1, j + 1, d)) return ORC_E_NOMEM;
        }
    return ORC_OK;
}

/* Wind at a particle: uniform, or bilinear on a regular grid (clamped at
 * the grid edge).  The reference hard-codes a uniform field
 * (wind_forcing.f90:36-37); the grid form carries the synthetic forcing
 * fields of the larger workloads. */
static void wind_at(const orc_params *prm, double px, double py, double *u,
                    double *v) {
    if (prm->wind_mode != 1) {
        *u = prm->wind[0];
        *v = prm->wind[1];
        return;
    }
    int nx = prm->grid_nx, ny = prm->grid_ny;
    double fx = (px - prm->grid_x0) / prm->grid_dx;
    double fy = (py - prm->grid_y0) / prm->grid_dy;
    double flx = floor(fx), fly = floor(fy);
    int i0 = (int)fmin(fmax(flx, 0.0), (double)(nx - 2));
    int j0 = (int)fmin(fmax(fly, 0.0), (double)(ny - 2));
    double tx = fmin(fmax(fx - (double)i0, 0.0), 1.0);
    double ty = fmin(fmax(fy - (double)j0, 0.0), 1.0);
    const double *U = prm->grid_u, *V = prm->grid_v;
    long r0 = (long)j0 * nx, r1 = r0 + nx;
    double u0 = (1.0 - tx) * U[r0 + i0] + tx * U[r0 + i0 + 1];
    double u1 = (1.0 - tx) * U[r1 + i0] + tx * U[r1 + i0 + 1];
    double v0 = (1.0 - tx) * V[r0 + i0] + tx * V[r0 + i0 + 1];
    double v1 = (1.0 - tx) * V[r1 + i0] + tx * V[r1 + i0 + 1];
    *u = (1.0 - ty) * u0 + ty * u1;
    *v = (1.0 - ty) * v0 + ty * v1;
}

/* The wind the forcing term sees at n points (x interleaved; uv receives
 * (u, v) interleaved): the sampler wind_forcing's term uses. */
void orc_wind_at(const orc_params *prm, int n, const double *x, double *uv) {
    for (int i = 0; i < n; ++i)
        wind_at(prm, x[2 * i], x[2 * i + 1], uv + 2 * i, uv + 2 * i + 1);
}

static int ensure_scratch(orc_ctx *c, int n) {
    if (n > c->ndr) {   /* zero-initialised like the module array */
        double *d = (double *)realloc(c->drp, (size_t)n * sizeof(double));
        if (!d) return -1;
        memset(d + c->ndr, 0, (size_t)(n - c->ndr) * sizeof(double));
        c->drp = d;
        c->ndr = n;
    }
    if (n <= c->nscr) return 0;
    free(c->div); free(c->grad); free(c->aux);
    c->div = (double *)malloc((size_t)n * sizeof(double));
    c->grad = (double *)malloc((size_t)n * 4 * sizeof(double));
    c->aux = (double *)malloc((size_t)n * 5 * sizeof(double));
    if (!c->div || !c->grad || !c->aux) { c->nscr = 0; return -1; }
    c->nscr = n;
    return 0;
}

/* ---- single_step, single_step.f90:5-108 ------------------------------ */
/* scale_h of each kernel (initialize.f90:208-217): 1 for the Wendland
 * kernels, 2 for the quadratic, 3 for the Gaussian, quintic and Marquis
 * kernels.  The search radius stays mhsml; q runs up to scale_h. */
static int scale_h(int skf) {
    if (skf == 1) return 2;
    if (skf == 2 || skf == 3 || skf == 7 || skf == 8) return 3;
    return 1;
}

/* One pair's kernel (kernel.f90): kernel_zero(cp_i), w(int_id) and the
 * gradient factor g with dwdx = g (dx / r) scale_h.  Fortran's x**k is
 * expanded as repeated products (ipow), default-REAL literals are exact
 * unless noted. */
static void kernel_pair(int skf, double r, double hij, double *kz, double *w, double *g) {
    const int sh = scale_h(skf);
    double q = r / hij * sh;
    double ih = 1.0 / hij;
    if (skf == 1) {
        /* Quadratic (kernel.f90:269-278) */
        double t = sh / hij;
        double factor = 2. / PI_ * (t * t);
        *kz = factor * 3. / 4.;
        *w = factor * (3. / 16. * ipow(q, 2) - 3. / 4. * q + 3. / 4.);
        *g = factor * (3. / 8. * q - 3. / 4.) / hij;
        return;
    }
    if (skf == 2) {
        /* Gaussian (kernel.f90:136-145) */
        double t = sh / hij;
        double factor = 1. / PI_ * (t * t);
        *kz = factor;
        *w = factor * exp(-ipow(q, 2));
        *g = factor * (-(2 * q * exp(-ipow(q, 2)))) / hij;
        return;
    }
    if (skf == 3) {
        /* quintic_spline (kernel.f90:193-221), three pieces of q */
        double t = sh / hij;
        double factor = 7.e0 / (478.e0 * PI_) * (t * t);
        *kz = factor * 66;        /* (3)**5 - 6 * (2)**5 + 15 */
        if (q >= 0. && q <= 1.) {
            *w = factor * (ipow(3 - q, 5) - 6 * ipow(2 - q, 5) + 15 * ipow(1 - q, 5));
            *g = factor * (-(5 * ipow(3 - q, 4)) + 30 * ipow(2 - q, 4) -
                           75 * ipow(1 - q, 4)) / hij;
        } else if (q > 1. && q <= 2.) {
            *w = factor * (ipow(3 - q, 5) - 6 * ipow(2 - q, 5));
            *g = factor * (-(5 * ipow(3 - q, 4)) + 30 * ipow(2 - q, 4)) / hij;
        } else {
            *w = factor * ipow(3 - q, 5);
            *g = factor * (-(5 * ipow(3 - q, 4))) / hij;
        }
        return;
    }
    if (skf == 7) {
        /* Marquis (kernel.f90:555-564): a smoothed step at q = 1 */
        double t = sh / hij;
        double factor = 1. / PI_ * (t * t);
        *kz = factor * (0.5 - 0.5 * tanh(0.));
        *w = factor * (0.5 - 0.5 * tanh(300 * (q - 1)));
        double ch = cosh(300 * (-1 + q));
        *g = factor * (-(150 * (1.0 / (ch * ch)))) / hij;
        return;
    }
    if (skf == 8) {
        /* Marquis2 (kernel.f90:604-621): C1 = 10, C2 = 50; 34.1558 and
         * log(2.) are default-REAL */
        const double C1 = 10., C2 = 50.;
        const double ln2 = (double)logf(2.0f);
        double factor = 1. / ((double)34.1558f * PI_ * ipow(hij, 2));
        *kz = factor * C2 * (log(cosh(C1)) / C1 + 1 + ln2 / C1);
        *w = factor * C2 * (log(cosh(C1 * (1 - q))) / C1 - q + 1 + ln2 / C1);
        *g = factor * (-(C2 * tanh(C1 * (1 - q))) - 1) / hij;
        return;
    }
    if (skf == 4) {
        /* Wendland_C2 (kernel.f90:387-396) */
        double factor = 3. / (2. * PI_) * (ih * ih);
        *kz = factor;
        *w = factor * (ipow(1. - q, 4) * (1 + 4 * q));
        *g = factor * (20. * (ipow(q, 4) - 3 * ipow(q, 3) + 3 * ipow(q, 2) - q)) / hij;
    } else if (skf == 5) {
        /* Wendland_C4 (kernel.f90:443-452) */
        double factor = 3. / PI_ * (ih * ih);
        *kz = factor * 3;
        *w = factor * (ipow(1 - q, 6) * (35 * ipow(q, 2) + 18 * q + 3));
        *g = factor * (-(56 * q * (5 * q + 1) * ipow(1 - q, 5))) / hij;
    } else {
        /* Wendland_C6 (kernel.f90:499-508) */
        double factor = 78. / (7. * PI_) * (ih * ih);
        double omq = 1 - q;
        *kz = factor;
        *w = factor * ipow(omq, 8) * (32 * ipow(q, 3) + 25 * ipow(q, 2) + 8 * q + 1);
        *g = factor * (-(22. * q * ipow(omq, 7) * (16 * ipow(q, 2) + 7. * q + 1))) / hij;
    }
}

/* The boundary kernel's value bw of one boundary pair (boundary_kernel.f90):
 * b_Wendland_C6 reads hsml_ij at the boundary-pair index (:491), C2 and C4
 * use b_length (:343, :417). */
static double boundary_w(int skf, double br, double hk) {
    const int sh = scale_h(skf);
    double q = br / B_LENGTH * sh;
    if (skf == 1) {
        /* b_Quadratic (boundary_kernel.f90:106-110) */
        double t = sh / B_LENGTH;
        double factor = 2. / PI_ * (t * t);
        return factor * (3. / 16. * ipow(q, 2) - 3. / 4. * q + 3. / 4.);
    }
    if (skf == 2 || skf == 7 || skf == 8) {
        /* b_Gaussian (boundary_kernel.f90:177-181), also for both Marquis
         * kernels (:38-43) */
        double t = sh / B_LENGTH;
        double factor = 1. / PI_ * (t * t);
        return factor * exp(-ipow(q, 2));
    }
    if (skf == 3) {
        /* b_quintic_spline (boundary_kernel.f90:250-272) */
        double t = sh / B_LENGTH;
        double factor = 15. / (478. * PI_) * (t * t);
        if (q >= 0. && q <= 1.)
            return factor * (ipow(3 - q, 5) - 6 * ipow(2 - q, 5) + 15 * ipow(1 - q, 5));
        if (q > 1. && q <= 2.) return factor * (ipow(3 - q, 5) - 6 * ipow(2 - q, 5));
        return factor * ipow(3 - q, 5);
    }
    if (skf == 4) {
        double t = 1.0 / B_LENGTH;
        double factor = 3. / (2. * PI_) * (t * t);
        return factor * (ipow(1. - q, 4) * (1 + 4 * q));
    }
    if (skf == 5) {
        double t = 1.0 / B_LENGTH;
        double factor = 3. / PI_ * (t * t);
        return factor * (ipow(1 - q, 6) * (35 * ipow(q, 2) + 18 * q + 3));
    }
    double ih = 1.0 / hk;
    double factor = 78. / (7. * PI_) * (ih * ih);
    double omq = 1 - q;
    return factor * ipow(omq, 8) * (32 * ipow(q, 3) + 25 * ipow(q, 2) + 8 * q + 1);
}

/* MC_bulk_viscosity and MC_shear_viscosity (bulk_viscosity.f90:113-143,
 * shear_viscosity.f90:91-122) at one particle. */
/* The boundary kernels' gradient (boundary_kernel.f90, bdwdx of each
 * b_ routine): bdwdx = -(g (bdx / br) scale_h) with g = factor inner /
 * b_length returned here, inner the routine's bracket in its evaluation
 * order. */
static double boundary_g(int skf, double br, double hk) {
    const int sh = scale_h(skf);
    double q = br / B_LENGTH * sh;
    double t = (double)sh / B_LENGTH, factor, inner;
    if (skf == 1) {            /* b_Quadratic (:106-112) */
        factor = 2. / PI_ * (t * t);
        inner = 3. / 8. * q - 3. / 4.;
    } else if (skf == 2 || skf == 7 || skf == 8) {   /* b_Gaussian (:177-183) */
        factor = 1. / PI_ * (t * t);
        inner = -(2 * q * exp(-ipow(q, 2)));
    } else if (skf == 3) {     /* b_quintic_spline (:250-275) */
        factor = 15. / (478. * PI_) * (t * t);
        if (q >= 0. && q <= 1.)
            inner = -(5 * ipow(3 - q, 4)) + 30 * ipow(2 - q, 4) - 75 * ipow(1 - q, 4);
        else if (q > 1. && q <= 2.)
            inner = -(5 * ipow(3 - q, 4)) + 30 * ipow(2 - q, 4);
        else
            inner = -(5 * ipow(3 - q, 4));
    } else if (skf == 4) {     /* b_Wendland_C2 (:343-349) */
        factor = 3. / (2. * PI_) * (t * t);
        inner = 20. * (ipow(q, 4) - 3 * ipow(q, 3) + 3 * ipow(q, 2) - q);
    } else if (skf == 5) {     /* b_Wendland_C4 (:417-423) */
        factor = 3. / PI_ * (t * t);
        inner = -(56 * q * (5 * q + 1) * ipow(1 - q, 5));
    } else {                   /* b_Wendland_C6 (:490-496): hsml_ij(int_id) */
        double ih = 1.0 / hk;
        factor = 78. / (7. * PI_) * (ih * ih);
        inner = -(22. * q * ipow(1 - q, 7) * (16 * ipow(q, 2) + 7. * q + 1));
    }
    return factor * inner / B_LENGTH;
}

static void mc_viscosity(double e11, double e12, double e22, double P, double *z,
                         double *et) {
    double sq = sqrt((e11 - e22) * (e11 - e22) + 4 * (e12 * e12));
    double e1 = 0.5 * ((e11 + e22) + sq);
    double e2 = 0.5 * ((e11 + e22) - sq);
    double eI = fabs(e1 + e2);
    if (eI < ZERO_MC) eI = ZERO_MC;
    double zz = 0.5 * P * (1 + K_T) / eI;
    if (zz > ZETA_MAX * P) zz = ZETA_MAX * P;
    double eII = e1 - e2;
    if (fabs(eII) < ZERO_MC) eII = ZERO_MC;
    double ee = MU_MC * (0.5 * P * (1 + K_T) - zz * (e1 + e2)) / eII;
    if (ee > ETA_MAX * P) ee = ETA_MAX * P;
    *z = zz;
    *et = ee;
}

/* c_yang (sound_speed.f90:57-58): sqrt(Y (1 - nu) / (rho (1 + nu)
 * (1 - 2 nu))), Y = 1.D9, nu = 0.3 (a default-REAL literal,
 * data_variables.f90:159-160) */
static const double Y_MOD = 1.0e9;
static const double NU = (double)0.3f;
double orc_sound_speed(double rho) {
    return sqrt(Y_MOD * (1 - NU) / (rho * (1 + NU) * (1 - 2 * NU)));
}

/* The sound speed single_step sees: sound_speed computes it after every
 * density update (time_integration.f90:140, 230, 274, 355, 383; and
 * initialize.f90:199-200) only when a switch reads it (sound_speed.f90:21);
 * otherwise it stays at its initial zero (initialize.f90:34). */
static int needs_c(const orc_params *prm) {
    return prm->art_visc || prm->tse == 1 || prm->boundary_friction;
}

/* regularization (regularization.f90:5-36): XSPH when requested, then the
 * artificial viscosity art_visc_alg names -- which runs whenever
 * regularization does, also for XSPH alone (then with the zero sound speed
 * the switches leave, :29-35). */
static void regularization(orc_ctx *c, const orc_params *prm, int n, const double *x,
                           const double *vx, const double *mass, const double *rho,
                           const double *hsml, double *dxdt, double *dvxdt) {
    double *cs = c->aux, *lv = c->aux + n, *xs = c->aux + 2 * (size_t)n;
    const int wc = needs_c(prm);
    for (int i = 0; i < n; ++i) cs[i] = wc ? orc_sound_speed(rho[i]) : 0.0;
    if (prm->xsph) {
        /* XSPH (:69-93): the same dvx = v_i - v_j enters both particles' sums */
        for (int i = 0; i < 2 * n; ++i) xs[i] = 0.;
        for (int k = 0; k < c->nip; ++k) {
            int i = c->p_i[k] - 1, j = c->p_j[k] - 1;
            double dv0 = vx[2 * i] - vx[2 * j], dv1 = vx[2 * i + 1] - vx[2 * j + 1];
            double rm = 0.5 * (rho[i] + rho[j]);
            xs[2 * i] = xs[2 * i] + mass[j] * dv0 * c->w[k] / rm;
            xs[2 * i + 1] = xs[2 * i + 1] + mass[j] * dv1 * c->w[k] / rm;
            xs[2 * j] = xs[2 * j] + mass[i] * dv0 * c->w[k] / rm;
            xs[2 * j + 1] = xs[2 * j + 1] + mass[i] * dv1 * c->w[k] / rm;
        }
        for (int i = 0; i < 2 * n; ++i) dxdt[i] = dxdt[i] - CHI_XSPH * xs[i];
    }
    const int alg = prm->art_visc_alg ? prm->art_visc_alg : 1;
    /* pi_ij keeps its value from the last pair that set it: a pair that
     * does not set it takes the previous one's (0 before the first) */
    double pi_ij = 0.;
    if (alg == 1) {
        /* art_visc_HK (:141-211): div_v over the pairs, zero where >= 0 */
        for (int i = 0; i < n; ++i) lv[i] = 0.;
        for (int k = 0; k < c->nip; ++k) {
            int i = c->p_i[k] - 1, j = c->p_j[k] - 1;
            double dv0 = vx[2 * i] - vx[2 * j], dv1 = vx[2 * i + 1] - vx[2 * j + 1];
            double w1 = dv0 * c->dwdx[2 * k], w4 = dv1 * c->dwdx[2 * k + 1];
            lv[i] = lv[i] - mass[j] * (w1 + w4) / rho[i];
            lv[j] = lv[j] - mass[i] * (w1 + w4) / rho[j];
        }
        for (int i = 0; i < n; ++i) {
            double dv = lv[i] >= 0. ? 0. : lv[i];
            double t = fabs(dv) * hsml[i];
            lv[i] = (ALPHA_PI * hsml[i] * cs[i] * fabs(dv) + BETA_PI * (t * t)) * rho[i];
        }
        for (int k = 0; k < c->nip; ++k) {
            int i = c->p_i[k] - 1, j = c->p_j[k] - 1;
            double dx0 = x[2 * i] - x[2 * j], dx1 = x[2 * i + 1] - x[2 * j + 1];
            double dv0 = vx[2 * i] - vx[2 * j], dv1 = vx[2 * i + 1] - vx[2 * j + 1];
            double crit = dv0 * dx0 + dv1 * dx1;
            if (crit <= 0.) pi_ij = lv[i] / (rho[i] * rho[i]) + lv[j] / (rho[j] * rho[j]);
            dvxdt[2 * i] = dvxdt[2 * i] - mass[j] * pi_ij * c->dwdx[2 * k];
            dvxdt[2 * j] = dvxdt[2 * j] + mass[i] * pi_ij * c->dwdx[2 * k];
            dvxdt[2 * i + 1] = dvxdt[2 * i + 1] - mass[j] * pi_ij * c->dwdx[2 * k + 1];
            dvxdt[2 * j + 1] = dvxdt[2 * j + 1] + mass[i] * pi_ij * c->dwdx[2 * k + 1];
        }
    } else if (alg == 2) {
        /* art_visc_Mon (:261-306); s_cor = 0.01 hsml_ij, 0.01 a default-REAL
         * literal */
        for (int k = 0; k < c->nip; ++k) {
            int i = c->p_i[k] - 1, j = c->p_j[k] - 1;
            double dx0 = x[2 * i] - x[2 * j], dx1 = x[2 * i + 1] - x[2 * j + 1];
            double dv0 = vx[2 * i] - vx[2 * j], dv1 = vx[2 * i + 1] - vx[2 * j + 1];
            double v_r = dv0 * dx0 + dv1 * dx1;
            double hij = c->hsml_ij[k], rij = c->r_ij[k];
            double s_cor = (double)0.01f * hij;
            if (v_r <= 0.) {
                double iota = hij * v_r / (rij * rij + s_cor * s_cor);
                pi_ij = (BETA_PI * (iota * iota) - ALPHA_PI * iota * (cs[i] + cs[j]) / 2) /
                        ((rho[i] + rho[j]) / 2);
            }
            dvxdt[2 * i] = dvxdt[2 * i] - mass[j] * pi_ij * c->dwdx[2 * k];
            dvxdt[2 * j] = dvxdt[2 * j] + mass[i] * pi_ij * c->dwdx[2 * k];
            dvxdt[2 * i + 1] = dvxdt[2 * i + 1] - mass[j] * pi_ij * c->dwdx[2 * k + 1];
            dvxdt[2 * j + 1] = dvxdt[2 * j + 1] + mass[i] * pi_ij * c->dwdx[2 * k + 1];
        }
    }
}

int orc_single_step(orc_ctx *c, const orc_params *prm, int n, const double *x,
                    const double *vx, const double *mass, const double *rho,
                    const double *hsml, const double *h, const double *A,
                    int nb, const double *xb, const double *massb,
                    double current_time, double *dxdt, double *dvxdt,
                    double *dhdt, double *dAdt, double *eps3, double *zeta,
                    double *eta, double *p, double *sig3) {
    if (n < 1) return ORC_E_ARG;
    if (ensure_scratch(c, n)) return ORC_E_NOMEM;
    if (n > c->nkz) {   /* zero-initialised like the module array */
        double *kz = (double *)realloc(c->kz, (size_t)n * sizeof(double));
        if (!kz) return ORC_E_NOMEM;
        memset(kz + c->nkz, 0, (size_t)(n - c->nkz) * sizeof(double));
        c->kz = kz;
        c->nkz = n;
    }
    double *div = c->div, *grad = c->grad;   /* grad(4, n) */
    /* reset (:36-44) */
    for (int i = 0; i < n; ++i) {
        dxdt[2 * i] = dxdt[2 * i + 1] = 0.;
        dvxdt[2 * i] = dvxdt[2 * i + 1] = 0.;
        dAdt[i] = dhdt[i] = 0.;
        div[i] = 0.;
        grad[4 * i] = grad[4 * i + 1] = grad[4 * i + 2] = grad[4 * i + 3] = 0.;
    }
    /* NNPS (:47); boundary pairs only with a boundary force (nnps.f90:440) */
    const int bforce = prm->boundary_pressure || prm->boundary_friction;
    int rc = orc_nnps(c, n, x, hsml, bforce ? nb : 0, xb);
    if (rc) return rc;
    /* pair loop (:56-75): kernel (Wendland C6 / C2 / C4, kernel.f90:471-520,
     * :359-408, :415-464; scale_h = 1 for all three, initialize.f90:208-209),
     * vector_divergence (vector_divergence.f90:27-28),
     * vector_gradient (vector_gradient.f90:27-40) */
    const int skf = prm->skf ? prm->skf : 6;
    for (int k = 0; k < c->nip; ++k) {
        int i = c->p_i[k] - 1, j = c->p_j[k] - 1;
        double dx0 = x[2 * i] - x[2 * j], dx1 = x[2 * i + 1] - x[2 * j + 1];
        double dv0 = vx[2 * i] - vx[2 * j], dv1 = vx[2 * i + 1] - vx[2 * j + 1];
        double r = c->r_ij[k], hij = c->hsml_ij[k];
        double g;
        kernel_pair(skf, r, hij, &c->kz[i], &c->w[k], &g);
        const double sh = scale_h(skf);
        double dw0 = g * (dx0 / r) * sh;
        double dw1 = g * (dx1 / r) * sh;
        c->dwdx[2 * k] = dw0;
        c->dwdx[2 * k + 1] = dw1;
        double w1 = dv0 * dw0, w2 = dv1 * dw0, w3 = dv0 * dw1, w4 = dv1 * dw1;
        div[i] = div[i] - mass[j] * (w1 + w4);
        div[j] = div[j] - mass[i] * (w1 + w4);
        grad[4 * i + 0] = grad[4 * i + 0] - mass[j] * w1 / rho[j];
        grad[4 * j + 0] = grad[4 * j + 0] - mass[i] * w1 / rho[i];
        grad[4 * i + 1] = grad[4 * i + 1] - mass[j] * w2 / rho[j];
        grad[4 * j + 1] = grad[4 * j + 1] - mass[i] * w2 / rho[i];
        grad[4 * i + 2] = grad[4 * i + 2] - mass[j] * w3 / rho[j];
        grad[4 * j + 2] = grad[4 * j + 2] - mass[i] * w3 / rho[i];
        grad[4 * i + 3] = grad[4 * i + 3] - mass[j] * w4 / rho[j];
        grad[4 * j + 3] = grad[4 * j + 3] - mass[i] * w4 / rho[i];
    }
    /* strain rate (:80-84), set only IF (pressure_force .OR. viscosity);
     * otherwise epsilon keeps its value, the zero initialize gives it
     * (initialize.f90:61-63) -- the switches are PARAMETERs, so it never
     * changes */
    const int strain = prm->pressure_force || prm->viscosity;
    for (int i = 0; i < n; ++i) {
        eps3[3 * i + 0] = strain ? grad[4 * i + 0] : 0.0;
        eps3[3 * i + 2] = strain ? grad[4 * i + 3] : 0.0;
        eps3[3 * i + 1] = strain ? 0.5 * (grad[4 * i + 1] + grad[4 * i + 2]) : 0.0;
    }
    /* internal_stress (internal_stress.f90:21-53): Hibler P (:24),
     * elliptical_bulk_viscosity (bulk_viscosity.f90:65-76),
     * elliptical_shear_viscosity (shear_viscosity.f90:57),
     * kreyscher_eos (eos.f90:141-147), stress_tensor (stress_tensor.f90:31-39) */
    const double em2 = 1.0 / (E_ELL * E_ELL);       /* e**(-2) */
    for (int i = 0; i < n; ++i) {
        double e11 = eps3[3 * i], e12 = eps3[3 * i + 1], e22 = eps3[3 * i + 2];
        double P = P_STAR * h[i] * exp(-(C_EOS * (1 - A[i])));
        double delta = sqrt((e11 * e11 + e22 * e22) * (1 + em2) +
                            4 * em2 * (e12 * e12) +
                            2 * e11 * e22 * (1 - em2));
        double dcl = fabs(delta) < DELTA_MIN ? DELTA_MIN : delta;
        double z = 0.5 * P * (1 + K_T) / dcl;
        double et = z / (E_ELL * E_ELL);
        if (prm->rheology == 2) mc_viscosity(e11, e12, e22, P, &z, &et);
        if (!prm->viscosity) {
            /* IF (viscosity) (internal_stress.f90:20-31) is skipped: zeta and
             * eta keep initialize's zero (initialize.f90:64-65), p its zero
             * (:30) unless eos sets it */
            z = 0.0;
            et = 0.0;
        }
        double pr = z * 2 * delta / (1 + K_T);               /* kreyscher_eos */
        if (!prm->pressure_force) {
            /* no eos call: p keeps the ice strength the viscosity block set
             * (internal_stress.f90:24), or zero without it */
            pr = prm->viscosity ? P : 0.0;
        } else if (prm->eos_alg == 1) {
            pr = P;                                           /* hibler_eos (eos.f90:59) */
        } else if (prm->eos_alg == 2) {
            /* gutfraind_eos (eos.f90:94-102): rho_max = 934, rho_0 = rho_ice,
             * negative pressures set to zero */
            double gam = 1.9e5 * h[i] * exp(-(20.0 * (1 - A[i])));
            pr = gam * (rho[i] - RHO_ICE) / (934.0 - rho[i]);
            if (pr < 0.) pr = 0.;
        }
        zeta[i] = z;
        eta[i] = et;
        p[i] = pr;
        sig3[3 * i + 0] = (z + et) * e11 + (z - et) * e22 - pr * (1 - K_T) / 2;
        sig3[3 * i + 1] = 2 * et * e12;
        sig3[3 * i + 2] = (z + et) * e22 + (z - et) * e11 - pr * (1 - K_T) / 2;
    }
    /* stress_2 (internal_stress.f90:189-216): (sigma_i + sigma_j) / (rho_i rho_j) */
    for (int k = 0; prm->pa_sph == 2 && k < c->nip; ++k) {
        int i = c->p_i[k] - 1, j = c->p_j[k] - 1;
        double dw0 = c->dwdx[2 * k], dw1 = c->dwdx[2 * k + 1];
        double s11i = sig3[3 * i], s12i = sig3[3 * i + 1], s22i = sig3[3 * i + 2];
        double s11j = sig3[3 * j], s12j = sig3[3 * j + 1], s22j = sig3[3 * j + 2];
        double cj = s11j * dw0 + s12j * dw1;
        double ci = s11i * dw0 + s12i * dw1;
        dvxdt[2 * i] = dvxdt[2 * i] + mass[j] / RHO_ICE / h[i] / rho[j] * (ci + cj);
        dvxdt[2 * j] = dvxdt[2 * j] - mass[i] / RHO_ICE / h[j] / rho[i] * (ci + cj);
        cj = s12j * dw0 + s22j * dw1;
        ci = s12i * dw0 + s22i * dw1;
        dvxdt[2 * i + 1] = dvxdt[2 * i + 1] + mass[j] / RHO_ICE / h[i] / rho[j] * (ci + cj);
        dvxdt[2 * j + 1] = dvxdt[2 * j + 1] - mass[i] / RHO_ICE / h[j] / rho[i] * (ci + cj);
    }
    /* stress_1 (internal_stress.f90:103-136) */
    for (int k = 0; prm->pa_sph != 2 && k < c->nip; ++k) {
        int i = c->p_i[k] - 1, j = c->p_j[k] - 1;
        double dw0 = c->dwdx[2 * k], dw1 = c->dwdx[2 * k + 1];
        double s11i = sig3[3 * i], s12i = sig3[3 * i + 1], s22i = sig3[3 * i + 2];
        double s11j = sig3[3 * j], s12j = sig3[3 * j + 1], s22j = sig3[3 * j + 2];
        double cj = (s11j * dw0 + s12j * dw1) / (rho[j] * rho[j]);
        double ci = (s11i * dw0 + s12i * dw1) / (rho[i] * rho[i]);
        dvxdt[2 * i] = dvxdt[2 * i] + rho[i] * mass[j] / RHO_ICE / h[i] * (ci + cj);
        dvxdt[2 * j] = dvxdt[2 * j] + rho[j] * mass[i] / RHO_ICE / h[j] * (-ci - cj);
        cj = (s12j * dw0 + s22j * dw1) / (rho[j] * rho[j]);
        ci = (s12i * dw0 + s22i * dw1) / (rho[i] * rho[i]);
        dvxdt[2 * i + 1] = dvxdt[2 * i + 1] + rho[i] * mass[j] / RHO_ICE / h[i] * (ci + cj);
        dvxdt[2 * j + 1] = dvxdt[2 * j + 1] + rho[j] * mass[i] / RHO_ICE / h[j] * (-ci - cj);
    }
    /* boundary_kernel -> b_Wendland_C6 / C2 / C4 (boundary_kernel.f90:438-504,
     * :290-357, :364-431) -> bf_Monaghan (boundary_force.f90:78-82).  The C6
     * kernel factor reads hsml_ij at the boundary-pair index, i.e. an entry
     * of the particle-pair list (:491); stale entries beyond nip are read as
     * they are. */
    const int bfa = prm->bfa ? prm->bfa : 1;
    const int bp = prm->boundary_pressure, bf = prm->boundary_friction;
    for (int k = 0; k < c->bnip; ++k) {
        int i = c->bp_j[k] - 1, b = c->bp_i[k] - 1;
        double br = c->br_ij[k];
        if (bfa == 3) {
            /* bf_Marquis (:210-221): nothing with the boundary pressure on;
             * the friction brakes the particle, less with distance */
            if (bp || !bf) continue;
            dvxdt[2 * i] = dvxdt[2 * i] - vx[2 * i] * (1 - (br / B_LENGTH));
            dvxdt[2 * i + 1] = dvxdt[2 * i + 1] - vx[2 * i + 1] * (1 - (br / B_LENGTH));
            continue;
        }
        if (bfa == 2) {
            /* bf_Wang (boundary_force.f90:145-174): the normal velocity and
             * the rest of the speed, each over dt; with friction the
             * tangential part is capped at xi times the normal force */
            if (!bp) continue;
            const double n0 = prm->b_normal ? prm->b_normal[2 * b] : 0.0;
            const double n1 = prm->b_normal ? prm->b_normal[2 * b + 1] : 0.0;
            double v0 = n0 * vx[2 * i] + n1 * vx[2 * i + 1];
            double v1 = sqrt(vx[2 * i] * vx[2 * i] + vx[2 * i + 1] * vx[2 * i + 1]) - v0;
            double F0 = mass[i] * v0 / prm->step_dt, F1 = mass[i] * v1 / prm->step_dt;
            if (bf && fabs(v1) > fabs(XI_FRIC * v0)) {
                v1 = v1 * XI_FRIC * fabs(F0 / F1);
                F1 = mass[i] * v1 / prm->step_dt;
            }
            dvxdt[2 * i] = dvxdt[2 * i] + (F0 * n0 + F1 * n1);
            dvxdt[2 * i + 1] = dvxdt[2 * i + 1] + (F0 * n1 - F1 * n0);
            continue;
        }
        double bdx0 = xb[2 * b] - x[2 * i], bdx1 = xb[2 * b + 1] - x[2 * i + 1];
        double hk = k < c->cap ? c->hsml_ij[k] : 0.0;
        if (bp) {
            double bw = boundary_w(skf, br, hk);
            double den = (mass[i] + massb[b]) * (br * br);
            double F0 = KAPPA_N * bdx0 * bw * 2 * massb[b] / den;
            double F1 = KAPPA_N * bdx1 * bw * 2 * massb[b] / den;
            dvxdt[2 * i] = dvxdt[2 * i] + (-F0) / rho[i];
            dvxdt[2 * i + 1] = dvxdt[2 * i + 1] + (-F1) / rho[i];
        }
        if (bf) {
            /* bf_Monaghan's friction (:87-100): kappa_t c v . bdx over br rho
             * against the boundary kernel's gradient */
            const double g = boundary_g(skf, br, hk);
            const double sh = scale_h(skf);
            const double bw0 = -(g * (bdx0 / br) * sh), bw1 = -(g * (bdx1 / br) * sh);
            const double ci = orc_sound_speed(rho[i]);
            const double v_r = vx[2 * i] * bdx0 + vx[2 * i + 1] * bdx1;
            const double pi_aj = KAPPA_T * ci * v_r / br / rho[i];
            dvxdt[2 * i] = dvxdt[2 * i] + (-(massb[b] * pi_aj * bw0)) / rho[i];
            dvxdt[2 * i + 1] = dvxdt[2 * i + 1] + (-(massb[b] * pi_aj * bw1)) / rho[i];
        }
    }
    /* wind_forcing (wind_forcing.f90:35-50) and water_drag
     * (water_drag.f90:35-51): both gated on current_time > 0 */
    if (prm->wind_force && current_time > 0.) {
        for (int i = 0; i < n; ++i) {
            double wu, wv;
            wind_at(prm, x[2 * i], x[2 * i + 1], &wu, &wv);
            double du = wu - vx[2 * i], dv = wv - vx[2 * i + 1];
            double hr = h[i] * RHO_ICE;
            dvxdt[2 * i] = dvxdt[2 * i] + (C_WIND * RHO_AIR * du * fabs(du)) / hr;
            dvxdt[2 * i + 1] = dvxdt[2 * i + 1] + (C_WIND * RHO_AIR * dv * fabs(dv)) / hr;
        }
    }
    if (prm->water_force && current_time > 0.) {
        for (int i = 0; i < n; ++i) {
            double du = prm->current[0] - vx[2 * i], dv = prm->current[1] - vx[2 * i + 1];
            double hr = h[i] * RHO_ICE;
            dvxdt[2 * i] = dvxdt[2 * i] + (C_WATER * RHO_WATER * du * fabs(du)) / hr;
            dvxdt[2 * i + 1] = dvxdt[2 * i + 1] + (C_WATER * RHO_WATER * dv * fabs(dv)) / hr;
        }
    }
    /* regularization (single_step.f90:99-101) */
    if (prm->xsph || prm->art_visc) regularization(c, prm, n, x, vx, mass, rho, hsml, dxdt, dvxdt);
    /* tendencies (:104-106) */
    for (int i = 0; i < n; ++i) {
        dhdt[i] = -(h[i] * div[i] / rho[i]);
        dAdt[i] = -(A[i] * div[i] / rho[i]);
        dxdt[2 * i] = dxdt[2 * i] + vx[2 * i];
        dxdt[2 * i + 1] = dxdt[2 * i + 1] + vx[2 * i + 1];
    }
    return ORC_OK;
}

/* ---- sum_density, density.f90:41-100 ---------------------------------- */
int orc_sum_density(orc_ctx *c, int n, const double *mass, double *rho) {
    if (n < 1 || n > c->nkz) return ORC_E_ARG;
    double *num = c->div, *den = c->grad;   /* scratch, n each */
    for (int i = 0; i < n; ++i) {
        num[i] = c->kz[i] * mass[i];
        den[i] = c->kz[i] * mass[i] / rho[i];
    }
    for (int k = 0; k < c->nip; ++k) {
        int i = c->p_i[k] - 1, j = c->p_j[k] - 1;
        den[i] = den[i] + mass[j] * c->w[k] / rho[j];
        den[j] = den[j] + mass[i] * c->w[k] / rho[i];
        num[i] = num[i] + mass[j] * c->w[k];
        num[j] = num[j] + mass[i] * c->w[k];
    }
    for (int i = 0; i < n; ++i) rho[i] = num[i] / den[i];
    return ORC_OK;
}

/* ---- con_density, density.f90:107-165 ---------------------------------
 * drhodt(i) = sum over the pairs of the last single_step, in list order,
 * of mass(j) (vx_i - vx_j) . dwdx_k (and mass(i) times the same for j),
 * with the velocities given (the module's at the time density() is
 * called) and the gradients the last single_step left. */
static int pair_rate(orc_ctx *c, int n, const double *vx, const double *mass,
                     double *drhodt) {
    if (n < 1) return ORC_E_ARG;
    for (int k = 0; k < c->nip; ++k) {
        int i = c->p_i[k] - 1, j = c->p_j[k] - 1;
        double dv0 = vx[2 * i] - vx[2 * j], dv1 = vx[2 * i + 1] - vx[2 * j + 1];
        double w1 = dv0 * c->dwdx[2 * k], w4 = dv1 * c->dwdx[2 * k + 1];
        double v_ij = w1 + w4;
        drhodt[i] = drhodt[i] + mass[j] * v_ij;
        drhodt[j] = drhodt[j] + mass[i] * v_ij;
    }
    return ORC_OK;
}

static int con_density(orc_ctx *c, int n, const double *vx, const double *mass,
                       double *drhodt) {
    if (n < 1) return ORC_E_ARG;
    for (int i = 0; i < n; ++i) drhodt[i] = 0.;
    return pair_rate(c, n, vx, mass, drhodt);
}

int orc_con_density(orc_ctx *c, int n, const double *vx, const double *mass,
                    double *drhodt) {
    return con_density(c, n, vx, mass, drhodt);
}

/* dwdx(2, nip) of the last single_step (kernel.f90, the kernel routines) */
void orc_get_dwdx(const orc_ctx *c, double *dwdx) {
    memcpy(dwdx, c->dwdx, sizeof(double) * 2 * (size_t)c->nip);
}

/* VP_time_step (time_step.f90:155) */
double orc_vp_dt(int n, const double *hsml) {
    double dt = 0.0;
    for (int i = 0; i < n; ++i) {
        double v = RHO_ICE * (hsml[i] * hsml[i]) * DELTA_MIN / (P_STAR * (1 + K_T));
        if (i == 0 || v < dt) dt = v;
    }
    return dt;
}

/* Monaghan_time_step (time_step.f90:39-125) without artificial viscosity
 * (artificial_visc = .FALSE., globals.f90:186): fF MINVAL(hsml / |a|)**0.5
 * with fF = 0.25 and a the last evaluation's dvxdt (:79-83), against
 * MINVAL(hsml / c) (:117), the smaller of the two (:121); c from c_yang
 * (sound_speed.f90), sqrt(Y (1 - nu) / (rho (1 + nu) (1 - 2 nu))), at the
 * step's final density.  Y = 1.D9, nu = 0.3 (a default-REAL literal,
 * data_variables.f90:159-160). */
/* Monaghan_time_step with artificial viscosity (time_step.f90:79-121): the
 * force condition as above, and the CFL condition with the velocity
 * divergence of the last single_step's pairs and gradients at the current
 * velocities and densities (:89-107, not clamped), 0.3 and 1.2 default-REAL
 * literals */
static double monaghan_dt_av(const orc_ctx *c, int n, const double *hsml, const double *rho,
                             const double *vx, const double *mass, const double *dvxdt,
                             double *div_v) {
    double mf = 0.0;
    for (int i = 0; i < n; ++i) {
        double acc = sqrt(dvxdt[2 * i] * dvxdt[2 * i] + dvxdt[2 * i + 1] * dvxdt[2 * i + 1]);
        double f = hsml[i] / acc;
        if (i == 0 || f < mf) mf = f;
        div_v[i] = 0.;
    }
    for (int k = 0; k < c->nip; ++k) {
        int i = c->p_i[k] - 1, j = c->p_j[k] - 1;
        double dv0 = vx[2 * i] - vx[2 * j], dv1 = vx[2 * i + 1] - vx[2 * j + 1];
        double w1 = dv0 * c->dwdx[2 * k], w4 = dv1 * c->dwdx[2 * k + 1];
        div_v[i] = div_v[i] - mass[j] * (w1 + w4) / rho[i];
        div_v[j] = div_v[j] - mass[i] * (w1 + w4) / rho[j];
    }
    double mc = 0.0;
    for (int i = 0; i < n; ++i) {
        double cs = orc_sound_speed(rho[i]);
        double ad = fabs(div_v[i]);
        double g = (double)0.3f * hsml[i] /
                   (cs + hsml[i] * ad + (double)1.2f * (ALPHA_PI * cs + BETA_PI * hsml[i] * ad));
        if (i == 0 || g < mc) mc = g;
    }
    double dt_force = 0.25 * sqrt(mf);
    return dt_force < mc ? dt_force : mc;
}

double orc_monaghan_dt(int n, const double *hsml, const double *rho, const double *dvxdt) {
    double mf = 0.0, mc = 0.0;
    for (int i = 0; i < n; ++i) {
        double acc = sqrt(dvxdt[2 * i] * dvxdt[2 * i] + dvxdt[2 * i + 1] * dvxdt[2 * i + 1]);
        double c = sqrt(Y_MOD * (1 - NU) / (rho[i] * (1 + NU) * (1 - 2 * NU)));
        double f = hsml[i] / acc, g = hsml[i] / c;
        if (i == 0 || f < mf) mf = f;
        if (i == 0 || g < mc) mc = g;
    }
    double dt_force = 0.25 * sqrt(mf);
    return dt_force < mc ? dt_force : mc;
}

/* mass_update_from_area (mass_evolution.f90:40-71) for sle = 0 and 1:
 * mass = rho_ice A h pi (alpha hsml)^2 with alpha = 1. / (2 hsml_multiplicity),
 * a default-REAL quotient (data_variables.f90:169); sle = 2 keeps the mass
 * (:27-29).  Called after every density update (time_integration.f90:137,
 * 227, 271, 352, 380). */
static void mass_evolution(const orc_params *prm, int n, const double *A, const double *h,
                           const double *hsml, double *mass) {
    if (prm->sle == 2) return;
    const double alpha = (double)(1.0f / 6);
    for (int i = 0; i < n; ++i) {
        double t = alpha * hsml[i];
        double area = PI_ * (t * t);
        mass[i] = RHO_ICE * A[i] * h[i] * area;
    }
}

/* ---- time_integration + predictor_corrector -------------------------- */
int orc_run_pc(orc_ctx *c, const orc_params *prm, int nsteps, int n, double *x,
               double *vx, double *mass, double *rho, double *hsml,
               double *h, double *A, int nb, const double *xb,
               const double *massb, double *dt_io, double *time_io,
               int *itimestep, double *eps3_o, double *zeta_o, double *eta_o,
               double *p_o, double *sig3_o) {
    if (n < 1 || nsteps < 0) return ORC_E_ARG;
    size_t n2 = (size_t)n * 2;
    double *buf = (double *)malloc(sizeof(double) * (size_t)n * 31);
    if (!buf) return ORC_E_NOMEM;
    double *x1 = buf, *v1 = x1 + n2, *h1 = v1 + n2, *A1 = h1 + n;
    double *dxdt = A1 + n, *dvxdt = dxdt + n2, *dhdt = dvxdt + n2, *dAdt = dhdt + n;
    double *eps3 = dAdt + n, *sig3 = eps3 + 3 * (size_t)n;
    double *zeta = sig3 + 3 * (size_t)n, *eta = zeta + n, *pr = eta + n;
    double *rho1 = pr + n, *drhodt = rho1 + n;
    const int cden = prm->den_alg == 2;   /* con_density */
    double dt = *dt_io, t = *time_io;
    int it = *itimestep;
    int rc = ORC_OK;
    const int scheme = prm->int_algorithm ? prm->int_algorithm : 2;
    for (int s = 0; s < nsteps && rc == ORC_OK; ++s) {
        /* monitor current quantities (:197-201, :323-327) */
        memcpy(x1, x, n2 * sizeof(double));
        memcpy(v1, vx, n2 * sizeof(double));
        memcpy(h1, h, (size_t)n * sizeof(double));
        memcpy(A1, A, (size_t)n * sizeof(double));
        memcpy(rho1, rho, (size_t)n * sizeof(double));
        orc_params sp = *prm;   /* the module dt single_step sees (bf_Wang) */
        sp.step_dt = dt;
        rc = orc_single_step(c, &sp, n, x, vx, mass, rho, hsml, h, A, nb, xb,
                             massb, t, dxdt, dvxdt, dhdt, dAdt, eps3, zeta,
                             eta, pr, sig3);
        if (rc) break;
        if (scheme == 1) {
            /* euler_scheme (:114-146): full step, density, Marquis_hsml */
            for (size_t i = 0; i < n2; ++i) {
                x[i] = x[i] + dxdt[i] * dt;
                vx[i] = vx[i] + dvxdt[i] * dt;
            }
            for (int i = 0; i < n; ++i) {
                h[i] = h[i] + dhdt[i] * dt;
                A[i] = A[i] + dAdt[i] * dt;
                if (A[i] > 1) A[i] = 1;
                if (prm->den_alg == 3) rho[i] = h[i] * RHO_ICE;
            }
            if (prm->den_alg == 1 && (rc = orc_sum_density(c, n, mass, rho))) break;
            if (cden) {
                /* (:131-134) on the updated velocities */
                if ((rc = con_density(c, n, vx, mass, drhodt))) break;
                for (int i = 0; i < n; ++i) rho[i] = rho[i] + drhodt[i] * dt;
            }
            goto end_of_step;
        }
        /* predict half step (:207-215), density (:218) */
        for (size_t i = 0; i < n2; ++i) {
            x[i] = x[i] + dxdt[i] * dt / 2;
            vx[i] = vx[i] + dvxdt[i] * dt / 2;
        }
        for (int i = 0; i < n; ++i) {
            h[i] = h[i] + dhdt[i] * dt / 2;
            A[i] = A[i] + dAdt[i] * dt / 2;
            if (A[i] > 1) A[i] = 1;
            if (prm->den_alg == 3) rho[i] = h[i] * RHO_ICE;   /* diagnostic_density */
        }
        if (prm->den_alg == 1 && (rc = orc_sum_density(c, n, mass, rho))) break;
        if (cden) {
            /* (:221-224, :347-349): the half-step velocities, the step-start
             * gradients */
            if ((rc = con_density(c, n, vx, mass, drhodt))) break;
            for (int i = 0; i < n; ++i) rho[i] = rho[i] + drhodt[i] * dt / 2;
        }
        mass_evolution(prm, n, A, h, hsml, mass);   /* (:227, :352) */
        rc = orc_single_step(c, &sp, n, x, vx, mass, rho, hsml, h, A, nb, xb,
                             massb, t, dxdt, dvxdt, dhdt, dAdt, eps3, zeta,
                             eta, pr, sig3);
        if (rc) break;
        if (scheme == 3) {
            /* leapfrog_scheme (:361-372): velocity, thickness, position from
             * the step-start values and the half-step tendencies, density */
            for (size_t i = 0; i < n2; ++i) {
                vx[i] = v1[i] + dvxdt[i] * dt;
                x[i] = x1[i] + (v1[i] + dxdt[i]) * dt / 2;
            }
            for (int i = 0; i < n; ++i) {
                h[i] = h1[i] + dhdt[i] * dt;
                A[i] = A1[i] + dAdt[i] * dt;
                if (A[i] > 1) A[i] = 1;
                if (prm->den_alg == 3) rho[i] = h[i] * RHO_ICE;
            }
            if (prm->den_alg == 1 && (rc = orc_sum_density(c, n, mass, rho))) break;
            if (cden) {
                /* (:375-377): the new velocities, the half-step gradients */
                if ((rc = con_density(c, n, vx, mass, drhodt))) break;
                for (int i = 0; i < n; ++i) rho[i] = rho1[i] + drhodt[i] * dt;
            }
            goto end_of_step;
        }
        /* corrector (:236-262); with summation density, density() after
         * the half-step values (:244) already changes rho; with continuity
         * density it takes the half-step velocities (vx is not updated yet)
         * and the half-step gradients, rho_12 = rho_1 + drhodt dt/2
         * (:245-248) and rho = 2 rho_12 - rho_1 (:265-268; the drhodt of
         * the third density call is not used) */
        if (prm->den_alg == 1 && (rc = orc_sum_density(c, n, mass, rho))) break;
        if (cden && (rc = con_density(c, n, vx, mass, drhodt))) break;
        for (size_t i = 0; i < n2; ++i) {
            double x12 = x1[i] + dxdt[i] * dt / 2;
            double v12 = v1[i] + dvxdt[i] * dt / 2;
            x[i] = 2 * x12 - x1[i];
            vx[i] = 2 * v12 - v1[i];
        }
        for (int i = 0; i < n; ++i) {
            double h12 = h1[i] + dhdt[i] * dt / 2;
            double A12 = A1[i] + dAdt[i] * dt / 2;
            h[i] = 2 * h12 - h1[i];
            A[i] = 2 * A12 - A1[i];
            if (A[i] > 1) A[i] = 1;
            if (prm->den_alg == 3) rho[i] = h[i] * RHO_ICE;
            if (cden) {
                double rho12 = rho1[i] + drhodt[i] * dt / 2;
                rho[i] = 2 * rho12 - rho1[i];
            }
        }
        if (prm->den_alg == 1 && (rc = orc_sum_density(c, n, mass, rho))) break;
    end_of_step:
        /* the predictor-corrector's third density call (:262) runs on the
         * final velocities with the half step's masses; only Benz_hsml reads
         * its continuity rate */
        if (cden && scheme == 2 && prm->sle == 1 &&
            (rc = con_density(c, n, vx, mass, drhodt)))
            break;
        mass_evolution(prm, n, A, h, hsml, mass);   /* (:137, :271, :380) */
        if (prm->sle == 2) {
            for (int i = 0; i < n; ++i) {
                /* Marquis_hsml (smoothing_length.f90:136-142) */
                hsml[i] = sqrt(mass[i] / rho[i]) * HSML_MULT;
                if (hsml[i] > prm->max_hsml) hsml[i] = prm->max_hsml;
            }
        } else if (prm->sle == 1) {
            /* Benz_hsml (smoothing_length.f90:75-109): the last single_step's
             * pairs and gradients at the step's final velocities; without
             * continuity density the sum goes onto the persistent drhodt,
             * with it drhodt is the last density call's (time_integration.f90
             * :262 in the predictor-corrector: final velocities too) */
            double *dr = c->drp;
            if (cden) {
                dr = drhodt;   /* the last density call's (:131, :262, :375) */
            } else if ((rc = pair_rate(c, n, vx, mass, c->drp))) {
                break;
            }
            for (int i = 0; i < n; ++i) {
                double dh = -(hsml[i] * dr[i] / 2 / rho[i]);
                hsml[i] = hsml[i] + dh * dt;
                if (hsml[i] < 0) hsml[i] = hsml[i] - dh * dt;
            }
        }
        /* time_step (:280) */
        if (prm->tse == 2) dt = orc_vp_dt(n, hsml);
        else if (prm->tse == 1 && prm->art_visc)
            dt = monaghan_dt_av(c, n, hsml, rho, vx, mass, dvxdt, drhodt);
        else if (prm->tse == 1) dt = orc_monaghan_dt(n, hsml, rho, dvxdt);
        /* loop bookkeeping (time_integration.f90:68-71) */
        it += 1;
        t = t + dt;
    }
    if (rc == ORC_OK) {
        if (eps3_o) memcpy(eps3_o, eps3, 3 * (size_t)n * sizeof(double));
        if (sig3_o) memcpy(sig3_o, sig3, 3 * (size_t)n * sizeof(double));
        if (zeta_o) memcpy(zeta_o, zeta, (size_t)n * sizeof(double));
        if (eta_o) memcpy(eta_o, eta, (size_t)n * sizeof(double));
        if (p_o) memcpy(p_o, pr, (size_t)n * sizeof(double));
    }
    *dt_io = dt;
    *time_io = t;
    *itimestep = it;
    free(buf);
    return rc;
}
