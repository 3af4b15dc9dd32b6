// simp_engine.h -- host side of the packed multi-instance engine.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "simp_device.h"

namespace simp {

// One simulation's host input (reference Fortran layout: x(2, n)).
struct InstanceInput {
  int n = 0, nb = 0;
  const double *x = nullptr, *vx = nullptr, *mass = nullptr, *rho = nullptr;
  const double *hsml = nullptr, *hth = nullptr, *A = nullptr;
  const double *xb = nullptr, *massb = nullptr;
  double dt = 0.0, time = 0.0;
  int itimestep = 0;
  int nsteps = 0;        // steps to run before the instance freezes (0: no limit)
  int tse = 2;
  int wind_mode = 0;
  double wind[2] = {0.0, 0.0}, current[2] = {0.0, 0.0};
  double max_hsml = 0.0;
  int den_alg = 3;       // 3: diagnostic density; 1: summation density
  int pa_sph = 1;        // 1: stress_1 (sigma / rho^2); 2: stress_2 (over rho_i rho_j)
  int eos_alg = 3;       // 1: Hibler P; 2: Gutfraind; 3: Kreyscher replacement pressure
  int rheology = 1;      // 1: elliptical yield curve; 2: Mohr-Coulomb
  int skf = 6;           // 6: Wendland C6; 4: Wendland C2; 5: Wendland C4
  int sle = 2;           // 2: Marquis hsml; 1: Benz hsml; 0: hsml unchanged
  int xsph = 0;          // 1: XSPH regularization
  int art_visc = 0;      // 1: artificial viscosity
  int art_visc_alg = 1;  // 1: Hernquist-Katz; 2: Monaghan
  int bfa = 1;           // 1: Monaghan; 2: Wang; 3: Marquis boundary force
  const double *bnormal = nullptr;   // (2, nb) boundary normals (Wang); null: zeros
  // the logical scenario switches of globals.f90:178-187 (without a
  // boundary force no boundary pairs exist: the boundary is dropped)
  int boundary_friction = 0, boundary_pressure = 1, pressure_force = 1, viscosity = 1;
  int wind_force = 1, water_force = 1;
  // optional wind grid
  int g_nx = 0, g_ny = 0;
  double g_x0 = 0, g_y0 = 0, g_dx = 1, g_dy = 1;
  const double *g_u = nullptr, *g_v = nullptr;
};

struct Profiler {
  struct Rec {
    std::string name;
    // each kernel of the launch: its start and end (the dispatch's own
    // timestamps, hipExtLaunchKernelGGL)
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  };
  std::vector<Rec> recs;
};

// Tuning parameters of the step kernels; none of them changes a result:
// the neighbours whose loads pass A's exact test and pass B issue together,
// and the candidate lists' skin as a fraction of MINVAL(hsml).
struct Tuning {
  int passA_batch = 2;   // 1 .. 4
  int passB_batch = 0;   // 2 .. 4; 0: 3 (mixed), 4 (FP64)
  float skin = 0.01f;    // (0, 0.2]
};

// Device bytes one instance occupies (state, per-evaluation buffers, cells,
// boundary-pair store), for admission against an HBM budget.
size_t instance_footprint(const InstanceInput &in);

class Engine {
 public:
  explicit Engine(int precision);
  ~Engine();

  // Replace the whole content by these instances.  With keep_store and a
  // single instance, the hsml_ij store of the previous content's first
  // instance is carried over (the reference's module array persists across
  // single_step calls).
  int upload(const std::vector<InstanceInput> &ins, bool keep_store);
  // Add instances (ids returned) / remove instances by id: the packed
  // layout is rebuilt on the device, live state is copied device to device.
  int add(const std::vector<InstanceInput> &ins, std::vector<int> *ids);
  // A new state for the single instance already resident -- same particle
  // and boundary counts and switches, its cells and boundary-pair store
  // large enough -- copied in through pinned staging; the layout, the
  // hsml_ij store and kernel_zero stay (they persist between single_step
  // calls as the reference's module arrays do).  Returns 1 when a full
  // upload is needed instead.
  // grid_same: a wind grid in `in` is the one of the resident instance
  int reload(const InstanceInput &in, bool grid_same);
  // tendencies and diagnostics of the single instance after eval_single, in
  // reference order, in one batch of copies through pinned staging (any
  // pointer may be null)
  int download_single(double *dxdt, double *dvxdt, double *dhdt, double *dAdt, double *eps3,
                      double *zeta, double *eta, double *p, double *sig3);
  // ---- one instance laid out in caller-provided device memory: nothing is
  // allocated, on the device or on the host.  The memory holds, in this
  // order, the persistent arrays -- the hsml_ij store (q_cap doubles), its
  // pair slots (q_cap int2) and kernel_zero (kz_cap doubles, by particle
  // index) -- then the layout of `in` and the staging of the call's inputs
  // and results.  The persistent arrays keep their contents between calls:
  // clear_persist clears them; otherwise kernel_zero entries from kz_valid
  // up to n start at zero (particles the last call did not have).  -3 when
  // the memory is too small (nothing is touched then), -5 when the engine
  // holds a layout it allocated.
  static size_t single_bytes(const InstanceInput &in, long q_cap, long kz_cap);
  int single_in_memory(char *base, size_t bytes, long q_cap, long kz_cap,
                       const InstanceInput &in, bool clear_persist, long kz_valid);
  // after eval_single: the results in the caller's order and layout, through
  // the staging area (any pointer may be null)
  int single_in_memory_out(double *dxdt, double *dvxdt, double *dhdt, double *dAdt,
                           double *eps3, double *zeta, double *eta, double *p, double *sig3);
  bool in_caller_memory() const { return arena_; }
  int remove(const std::vector<int> &ids);
  // Remove instances by id and add new ones in one layout rebuild.
  int update(const std::vector<int> &remove_ids, const std::vector<InstanceInput> &ins,
             std::vector<int> *ids);
  int index_of(int id) const;
  const std::vector<int> &ids() const { return ids_; }

  // One right-hand-side evaluation of the state S (single_step); writes
  // tendencies and diagnostics on the device.
  int eval_single(bool want_diag);
  // nsteps predictor-corrector steps of every active instance.
  int steps(int nsteps, bool use_graph);
  int sync_and_check();
  // host copies (by packed index)
  int download_state(int inst, double *x, double *vx, double *rho, double *hsml, double *hth,
                     double *A);
  int download_scalars(int inst, double *dt, double *time, int *itimestep, int *nip, int *bnip,
                       int *err);
  int download_status(int inst, int *itimestep, int *steps_left, int *active, int *err);
  int download_all_status(std::vector<InstDyn> &out);
  // interaction statistics of instance k's last collecting evaluation
  // (-5: none yet): max, min, sum over particles, particles with none,
  // particle pairs, and the evaluation's itimestep
  int interaction_stats(int k, int *maxip, int *minip, long long *sumip, int *noip,
                        long long *nip, int *step);
  int download_tend(int inst, double *dxdt, double *dvxdt, double *dhdt, double *dAdt);
  int download_diag(int inst, double *eps3, double *zeta, double *eta, double *p,
                    double *sig3);
  int export_pairs(int cap, int *nip, int *pi, int *pj, double *rij, double *hij, int bcap,
                   int *bnip, int *bpi, int *bpj, double *brij, double *w = nullptr,
                   double *dwdx = nullptr);
  // kernel_zero of instance inst (n values)
  int download_kernel_zero(int inst, double *kz);
  // snapshots: the arrays of instance k a restart needs besides its state
  // -- masses (reference order), boundary particles, kernel_zero and the
  // hsml_ij store (store_cap entries) -- and their restore after an upload
  long store_cap(int inst) const;
  // (drp: Benz's persistent drhodt, by reference index; bnormal: (2, nb)
  // boundary normals of the Wang force)
  int download_persist(int inst, double *mass, double *xb, double *massb, double *kz,
                       double *hq, double *drp = nullptr, double *bnormal = nullptr);
  int upload_persist(int inst, const double *kz, const double *hq, long nq,
                     const double *drp = nullptr);
  int cur_set(int inst);

  // 1: rebuild the cell and candidate lists at every evaluation (the
  // reference searches anew every time; the results are the same)
  void set_force_search(bool on);
  // instance k's evaluations since its upload, and of them those that
  // rebuilt the cell lists / the candidate lists
  int search_stats(int k, long long *evals, long long *cell, long long *cand);
  void set_tuning(const Tuning &t);
  const Tuning &tuning() const { return tune_; }
  // The autotuner: times candidate settings over `steps` steps of the
  // resident content each and keeps the fastest -- pass A's batch, then pass
  // B's, then the skin, one coordinate at a time.  The trial steps are steps
  // of the simulation (the results do not depend on the setting).  us_start:
  // microseconds per step of the setting it started from; us_best: of the
  // one it keeps.
  int autotune(int nsteps, double *us_start, double *us_best);
  // time integration of steps(): 2 predictor_corrector, 1 euler_scheme,
  // 3 leapfrog_scheme (time_integration.f90:155-282, 86-148, 289-391)
  void set_scheme(int int_algorithm);
  void set_diag(bool on) { diag_ = on; }
  bool diag() const { return diag_; }
  hipStream_t stream() const { return st_; }
  int ninst() const { return (int)meta_.size(); }
  long total_particles() const { return d_.ntot; }
  const InstMeta &meta(int k) const { return meta_[k]; }
  Profiler *prof = nullptr;

 private:
  int repack(const std::vector<int> &keep, const std::vector<InstanceInput> &add,
             int carry_store_from, std::vector<int> *new_ids);
  void enqueue_eval(int buf, int phase, bool want_diag);
  void enqueue_prefix(int buf);
  void enqueue_step();
  int graph_build();
  void graph_drop();
  template <class F>
  void L(const char *name, F &&f);

  int stage_reserve(size_t bytes);

  int precision_;
  hipStream_t st_ = nullptr;
  std::vector<BlockEntry> blk_host_;  // the block table as uploaded (state set 0)
  void *stage_ = nullptr;             // pinned host staging (reload, download_single)
  size_t stage_bytes_ = 0;
  DevState d_{};
  std::vector<void *> allocs_;        // device buffers of d_
  int *scan_tmp_ = nullptr;
  std::vector<InstMeta> meta_;
  std::vector<int> ids_;
  int next_id_ = 0;
  bool diag_ = false;
  bool any_boundary_ = false;
  bool any_xsw_ = false;              // some instance uses a non-default switch value
  bool any_cden_ = false;             // some instance uses continuity density
  bool any_av_ = false;               // some instance runs an artificial viscosity
  bool any_endp_ = false;             // some instance needs the end-of-step pair pass
  hipGraphExec_t gexec_ = nullptr;
  hipGraph_t graph_ = nullptr;
  bool force_search_ = false;
  int scheme_ = 2;
  Tuning tune_;
  bool arena_ = false;                // d_ lies in caller memory (single_in_memory)
  double *arena_stage_ = nullptr;     // its staging area
};

}  // namespace simp
