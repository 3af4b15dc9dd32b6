// simp_device.h -- device data layout shared by the kernels and the host
// engine.  Particles of many independent simulations ("instances") are
// packed into one set of structure-of-arrays buffers; every instance owns a
// contiguous particle range, boundary range, cell range and boundary-pair
// range, described by InstMeta (fixed between uploads) and InstDyn (changes
// every evaluation).  A particle's slot in its instance's range is its
// position in the cell lists: whenever an instance's lists are rebuilt its
// particle state moves with them (from one of two state sets to the
// other), so the gathers read the state itself, in list order, and orig[]
// keeps each slot's reference particle index.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace simp {

// Physical constants of data_variables.f90:153-180 and globals.f90:199.
// rho_air is a default-REAL literal widened to REAL(8) in the reference.
constexpr double kPi = 3.1415926535897932;
constexpr double kPStar = 2.75e4;
constexpr double kCEos = 20.0;
constexpr double kRhoIce = 900.0;
constexpr double kRhoMaxEos = 934.0;   // gutfraind_eos rho_max (eos.f90:88)
constexpr double kRhoWater = 1026.0;
constexpr double kRhoAir = (double)1.3f;
constexpr double kCWind = 1.2e-3;
constexpr double kCWater = 5.5e-3;
constexpr double kDeltaMin = 2.0e-9;
constexpr double kBLength = 8000.0;
constexpr double kKappaN = 5.4e12;
constexpr double kKT = 0.0;
constexpr double kE = 2.0;
constexpr double kHsmlMult = 3.0;
// c_yang's material constants (sound_speed.f90; data_variables.f90:159-160):
// Young modulus and Poisson ratio, the latter a default-REAL literal
constexpr double kYoung = 1.0e9;
constexpr double kNu = (double)0.3f;
// artificial viscosity, XSPH (data_variables.f90:171-173): alpha_pi = 1.,
// beta_pi = 2. (exact), chi = 0.3 (default REAL)
constexpr double kAlphaPi = 1.0;
constexpr double kBetaPi = 2.0;
constexpr double kChi = (double)0.3f;
// boundary friction (data_variables.f90:175-177): kappa_t = 5.D-8, xi =
// 0.5225 (default REAL)
constexpr double kKappaT = 5.e-8;
constexpr double kXi = (double)0.5225f;
constexpr int kMaxIndInt = 200;     // globals.f90:24
constexpr int kPrintStep = 2000;    // globals.f90:55

constexpr int kTPB = 256;           // threads per block of particle kernels
constexpr int kScanItems = 2048;    // elements per block of the scan
constexpr int kMaxBoundaryPerInst = 8192;
constexpr int kNbrMax = 32;         // neighbours kept per particle in the list
constexpr int kPairBlk = 256;       // pair ranks per block of the viscosity's pi_ij passes
// per-block partial reductions: min/max x, min/max y, min hsml, max speed,
// max hsml change, and for the Monaghan time step min hsml/|a|, min hsml/c
constexpr int kParts = 9;

// Error bits (OR-ed into InstDyn::err); translated to SIMP_E_* on the host.
constexpr int kErrCapacity = 1;
constexpr int kErrNan = 2;
constexpr int kErrBndNbr = 4;
constexpr int kErrSync = 8;     // the rebuild kernel's grid barrier timed out
constexpr int kBarWords = 8 * 32 + 2;   // its arrival counters and generation

struct InstMeta {
  int p_off, n;          // particle range
  int blk_off, nblk;     // logical particle-block range
  int b_off, nb;         // boundary-particle range
  int c_off, c_cap;      // cell range (capacity)
  int q_off, q_cap;      // boundary-pair / hsml_ij-store range
  int tse, wind_mode;
  int g_off, g_nx, g_ny;  // wind grid range
  int den_alg;           // 3: rho = h rho_ice; 1: summation density
  int pa_sph;            // 1: stress_1; 2: stress_2
  int eos_alg;           // 1: Hibler P; 2: Gutfraind; 3: Kreyscher
  int rheology;          // 1: elliptical yield curve; 2: Mohr-Coulomb
  int skf;               // 6: Wendland C6; 4: C2; 5: C4; 1 quadratic, 2 Gaussian,
                         // 3 quintic spline, 7 / 8 Marquis
  int sle;               // 2: Marquis hsml; 1: Benz hsml; 0: hsml unchanged
  int xsph;              // 1: XSPH position regularization
  int avis;              // artificial viscosity regularization runs with: 0 none,
                         // 1 Hernquist-Katz, 2 Monaghan (whenever XSPH or
                         // artificial_visc is on, regularization.f90:29-35)
  int needc;             // 1: the sound speed is c_yang of the density (else 0)
  int avdt;              // 1: the Monaghan time step takes the viscous CFL terms
  int bfa;               // boundary force: 1 Monaghan, 2 Wang, 3 Marquis
  int endp;              // 1: an end-of-step pair pass (Benz hsml, viscous CFL)
  int fric;              // 1: boundary friction (boundary_friction)
  int bpres;             // 1: boundary normal force (boundary_pressure)
  int wind_on, water_on; // wind_force, water_force (globals.f90:182-183)
  int novisc;            // 1: viscosity off -- zeta, eta stay zero (internal_
                         //    stress.f90:20-31); eos_alg 0 then: p stays zero
  int nostrain;          // 1: viscosity and pressure_force off -- the strain rate
                         //    is never set and stays zero (single_step.f90:80)
  int pad_;
  double g_x0, g_y0, g_dx, g_dy;
  double wind_u, wind_v, cur_u, cur_v;
  double max_hsml;
  double bxmin, bxmax, bymin, bymax;   // boundary bounding box (static)
};

struct InstDyn {
  double dt, time;
  int itimestep, err;
  // grid of the current evaluation (bucket_search :211-224)
  double bin;
  int xlo, ylo, W, H;
  int nip, bnip;
  int stamp;            // evaluation counter (boundary cell flags)
  float pf_slack;       // bound of the prefilter's FP32 distance error (m)
  int active;           // 0: frozen (finished or not started), kernels skip it
  int steps_left;       // steps until the instance freezes; < 0: no limit
  int lstamp;           // stamp of the evaluation that built the boundary cell lists
  int rlim, rstop;      // k_ranks: rank limit, list positions whose pairs rank below it
  int gendI, gendX;     //   and the pair prefixes at the end of the lists
  int rchg;             // 1: the pair ranks below the limit must be found again
  int rfresh;           // 1: pass B takes the hsml_ij of the ranks below the limit
                        //    from their pair slots (and stores it)
  double hmin;          // MINVAL(hsml) of the current evaluation
  // candidate lists (kept from the evaluation that built them while they
  // still cover the pair set): built, their skin, the skin of lists built
  // in this evaluation (0: none), the bound of 2 max|r - r_built| +
  // max|hsml - hsml_built| for the state the next evaluation sees, the
  // largest speed of the last evaluated state
  int vbuilt;
  float vskin, vskin_next, vmoved, vprev;
  // interaction statistics (interaction_stats, nnps.f90:613-658) of the last
  // evaluation that collected them: every evaluation of a step that is a
  // multiple of print_step, and every single evaluation.  ipstep: its
  // itimestep (-1: none yet); per-particle interactions = pairs + boundary
  // pairs of the particle
  int ipstep, ipmax, ipmin, ipzero;
  unsigned long long ipsum, ipown;    // sum of interactions; particle pairs
  unsigned long long n_eval, n_cell, n_cand;   // evaluations; of them, those that
                        // rebuilt the cell lists / the candidate lists
  int sel;              // state set holding the instance at the start of an evaluation
  int rb;               // 1: its lists are rebuilt in this evaluation (the state
                        //    then moves to set sel ^ 1: the current set is sel ^ rb)
};

// Per-launch block table entry: one block covers kTPB particle slots of
// exactly one instance.
struct BlockEntry {
  int inst, first;   // instance, first local slot
  int lid;           // logical block id (index of the block's partials)
  int nloc;          // slots in the block (min(n - first, kTPB))
  int pos;           // packed index of the first slot (p_off + first)
  int cs;            // the instance's current state set (sel ^ rb): 0 after a
                     // repack, moved by k_rebuild when the instance rebuilds,
                     // unchanged by the fold that ends the evaluation
  int tmass;         // 1: the instance's masses evolve (sle 0 or 1) and the
                     // half-step state's are in masst
  int pad;
};

// Device pointers of one engine.
struct DevState {
  // particle state by slot (= cell-list position), in two sets: index
  // [2 * set + buf], buf 0 = S (step start / current), 1 = T (half step)
  double *x[4], *y[4], *vx[4], *vy[4], *h[4], *A[4], *rho[4], *hsml[4];
  double *mass[2];
  // the half-step state's masses of instances whose masses evolve (sle 0, 1:
  // mass_evolution.f90:40-71 after every density update), by set; null when
  // no instance's do
  double *masst[2];
  int *orig[2];         // reference (0-based) particle index of each slot
  int *cellid[2];       // cell (local id) of each slot
  // cell lists: members per cell (kept current by the key updates), their
  // prefix, and for a rebuild the slot of each particle in its new cell
  // (from fill, zero between rebuilds) and the new order (cl[new slot] =
  // old slot)
  int *count, *cell_start, *scan_tmp, *fill, *slot, *cl;
  int *cinst;           // instance of each cell
  // per cell: (stamp of the evaluation << 32) | largest member hsml as
  // FP32 bits, rounded up
  unsigned long long *cellh;
  int *rebuild;          // 1: the cell lists are rebuilt in this evaluation
  unsigned *bar;         // the rebuild kernel's grid barrier (kBarWords words)
  float4 *pf;            // FP32 (x, y, hsml) by slot, x and y relative to the
                         // grid origin: the candidate prefilter
  // neighbour list of the evaluation, built by pass A in the reference's
  // pair order for each particle: nbr[j * ntot + s] = list position of the
  // j-th neighbour of position s (kNbrMax + 1 rows, the last one spare for
  // pass A's stores past a full list); nnb[s] = count (> kNbrMax: not
  // stored, the gathers rescan the 3x3 block)
  int *nbr, *nnb;
  // candidate lists: per slot, the list positions of the 3x3 block within
  // mhsml + skin at the evaluation that built them (j-major like nbr, at
  // most kNbrMax; vnc > kNbrMax: not stored)
  int *vnb, *vnc;
  double *part;          // per-block partial reductions (kParts x nblk)
  int *bflag;            // per cell: stamp of the last evaluation in which a
                         // boundary particle's 3x3 block covered the cell
  int *own_cnt, *moff, *pair_count, *pair_start, *cell_intra;
  int2 *gpre;            // per list position: pair prefixes (intra, cross) of k_ranks
  int *osig;             // per list position below the ranks' stop: a hash of the
                         // slots of the particle's owned neighbours
  int2 *hpair;           // per rank below the limit: the pair's owner and partner
                         // slots (-1: no pair, the hsml_ij entry keeps its value)
  void *sig11, *sig12, *sig22, *div;   // pass-A outputs (FP64 variant: the stress)
  // mixed variant: pass A's record of each slot for pass B's stress sums,
  // FP32 {sig11, sig12, sig22} / rho^2 and the mass (in sig11's allocation)
  float4 *srec;
  // summation density (den_alg = 1): kernel_zero of each particle (by
  // p_off + reference index; persists across evaluations like the
  // reference's module array), and per slot the numerator and the density
  // that pass A's sums give for the state it evaluated
  double *kz, *dnum, *dsum;
  double *diag;          // eps11, eps12, eps22, zeta, eta, p (6 x N by slot), optional
  double *sig_out;       // sigma_11, sigma_12, sigma_22 (3 x N by slot), with diag
  double *tend;          // dxdt, dydt, dvxdt, dvydt, dhdt, dAdt (6 x N by slot), optional
  // boundary
  double *xb, *yb, *massb;
  int *bpre, *bbase, *bcount;
  int *bcnt;             // per (boundary particle, cell of its block): pair counts
  int *bkey;             // cell key of each boundary particle when the lists were built
  // per cell covered by a boundary particle's 3x3 block: the run
  // [brun.x, brun.y) of blist holding those boundary particles (local index
  // | slot of the cell in its block << 24) in ascending index; bcellc holds
  // the stamp of the evaluation whose builder last claimed the cell
  int2 *brun;
  int *blist, *bcellc;
  double *hq;            // persistent hsml_ij store (stale semantics)
  // switches beyond the defaults (allocated only when an instance uses one):
  double *bnx, *bny;     // boundary normals (the Wang boundary force)
  double *drp;           // Benz's drhodt, by p_off + reference index: persists
                         // across steps like the reference's module array
  double *lv;            // per slot: the Hernquist-Katz viscosity's lv
  int *pidx;             // per (list entry, slot), j-major like nbr: the pair's
                         // rank in the packed pair ranks (pair_start)
  double *piv;           // per pair rank: the artificial viscosity's pi_ij
  int *pflg;             // per pair rank: 1 where the pair sets pi_ij
  int *pblast;           // per block of kPairBlk ranks: its last setting rank (-1: none)
  double *hnew, *cfl;    // per slot, Euler: the end pass's hsml and CFL term
  int pcap;              // capacity of piv / pflg (pair ranks)
  double *grid;          // wind grids, u then v per instance
  InstMeta *meta;
  InstDyn *dyn;
  BlockEntry *blk;       // particle blocks
  int nblk;
  int ntot, nbtot, ctot, qtot, ninst;
  int force_search;      // 1: rebuild the cell and candidate lists at every evaluation
  float skin;            // candidate-list skin, a fraction of MINVAL(hsml) (tuned)
};

}  // namespace simp
