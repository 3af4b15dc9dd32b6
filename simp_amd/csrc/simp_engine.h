// simp_engine.h -- host side of the packed multi-instance engine.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
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
  int tse = 2;
  int wind_mode = 0;
  double wind[2] = {0.0, 0.0}, current[2] = {0.0, 0.0};
  double max_hsml = 0.0;
  // optional wind grid
  int g_nx = 0, g_ny = 0;
  double g_x0 = 0, g_y0 = 0, g_dx = 1, g_dy = 1;
  const double *g_u = nullptr, *g_v = nullptr;
  // capacity hints (0 = derive)
  int c_cap = 0, q_cap = 0;
};

struct Profiler {
  struct Rec {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
};

class Engine {
 public:
  explicit Engine(int precision);
  ~Engine();

  // Pack and upload the instances (replaces the previous content).  The
  // hsml_ij store keeps its contents when the layout is unchanged.
  int upload(const std::vector<InstanceInput> &ins, bool keep_store);
  // One right-hand-side evaluation of the state S (single_step); writes
  // tendencies and diagnostics on the device.
  int eval_single(bool want_diag);
  // nsteps predictor-corrector steps.
  int steps(int nsteps, bool use_graph);
  int sync_and_check();
  // host copies
  int download_state(int inst, double *x, double *vx, double *rho, double *hsml, double *hth,
                     double *A);
  int download_scalars(int inst, double *dt, double *time, int *itimestep, int *nip, int *bnip,
                       int *err);
  int download_tend(int inst, double *dxdt, double *dvxdt, double *dhdt, double *dAdt);
  int download_diag(int inst, double *eps3, double *zeta, double *eta, double *p,
                    double *sig3);
  int export_pairs(int cap, int *nip, int *pi, int *pj, double *rij, double *hij, int bcap,
                   int *bnip, int *bpi, int *bpj, double *brij);
  // A probe evaluation prefix (binning + boundary prep) to size the
  // boundary-pair store; returns bnip per instance.
  int probe_bnip(std::vector<int> &bnip);

  void set_diag(bool on) { diag_ = on; }
  bool diag() const { return diag_; }
  hipStream_t stream() const { return st_; }
  int ninst() const { return (int)meta_.size(); }
  const InstMeta &meta(int k) const { return meta_[k]; }
  Profiler *prof = nullptr;

 private:
  void enqueue_eval(int buf, int phase, bool want_diag);
  void enqueue_prefix(int buf, bool fork_boundary);
  void enqueue_step();
  int ensure(int ntot, int nbtot, int ctot, int qtot, int ninst, int nblk, long gtot);
  void free_all();
  int graph_build();
  template <class F>
  void L(const char *name, F &&f);

  int precision_;
  hipStream_t st_ = nullptr;
  hipStream_t side_ = nullptr;        // boundary prep, concurrent with pass A
  hipEvent_t ev_fork_ = nullptr, ev_join_ = nullptr;
  DevState d_{};
  std::vector<InstMeta> meta_;
  bool diag_ = false;
  bool any_boundary_ = false;
  // capacities of the current allocation
  int cap_n_ = 0, cap_nb_ = 0, cap_c_ = 0, cap_q_ = 0, cap_inst_ = 0, cap_blk_ = 0;
  long cap_g_ = 0;
  int *scan_tmp_ = nullptr;
  int cap_scan_ = 0;
  hipGraphExec_t gexec_ = nullptr;
  hipGraph_t graph_ = nullptr;
  std::vector<void *> allocs_;
};

}  // namespace simp
