// simp_engine.hip -- packed multi-instance engine: device buffers, uploads,
// the per-evaluation kernel pipeline, the captured step graph, downloads.
#include "simp_engine.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "simp_kernels.h"

namespace simp {

#define HIPCHK(x)                       \
  do {                                  \
    if ((x) != hipSuccess) return -2;   \
  } while (0)

static inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// Host copy of the bucket_search grid size (nnps.f90:211-224) for sizing.
static long grid_cells(int n, const double *x, const double *hsml, int nb, const double *xb,
                       double shrink) {
  double hmin = hsml[0];
  double xmn = x[0], xmx = x[0], ymn = x[1], ymx = x[1];
  for (int i = 0; i < n; ++i) {
    hmin = std::min(hmin, hsml[i]);
    xmn = std::min(xmn, x[2 * i]); xmx = std::max(xmx, x[2 * i]);
    ymn = std::min(ymn, x[2 * i + 1]); ymx = std::max(ymx, x[2 * i + 1]);
  }
  for (int b = 0; b < nb; ++b) {
    xmn = std::min(xmn, xb[2 * b]); xmx = std::max(xmx, xb[2 * b]);
    ymn = std::min(ymn, xb[2 * b + 1]); ymx = std::max(ymx, xb[2 * b + 1]);
  }
  if (!(hmin > 0) || !std::isfinite(xmn + xmx + ymn + ymx)) return -1;
  double bin = 2 * hmin * shrink;
  long W = (long)std::round(xmx / bin) - (long)std::round(xmn / bin) + 3;
  long H = (long)std::round(ymx / bin) - (long)std::round(ymn / bin) + 3;
  return W * H;
}

Engine::Engine(int precision) : precision_(precision) {
  hipStreamCreate(&st_);
  hipStreamCreate(&side_);
  hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming);
  hipEventCreateWithFlags(&ev_join_, hipEventDisableTiming);
}

Engine::~Engine() {
  if (gexec_) hipGraphExecDestroy(gexec_);
  if (graph_) hipGraphDestroy(graph_);
  free_all();
  if (d_.hq) hipFree(d_.hq);
  if (ev_fork_) hipEventDestroy(ev_fork_);
  if (ev_join_) hipEventDestroy(ev_join_);
  if (side_) hipStreamDestroy(side_);
  if (st_) hipStreamDestroy(st_);
}

void Engine::free_all() {
  for (void *p : allocs_) hipFree(p);
  allocs_.clear();
  if (scan_tmp_) hipFree(scan_tmp_);
  scan_tmp_ = nullptr;
  cap_n_ = cap_nb_ = cap_c_ = cap_inst_ = cap_blk_ = 0;
  cap_g_ = 0;
  cap_scan_ = 0;
  // the hsml_ij store is owned separately (kept across uploads)
}

template <class T>
static int dalloc(std::vector<void *> &list, T **p, size_t count) {
  if (count == 0) count = 1;
  if (hipMalloc((void **)p, count * sizeof(T)) != hipSuccess) return -6;
  list.push_back(*p);
  return 0;
}

int Engine::ensure(int ntot, int nbtot, int ctot, int qtot, int ninst, int nblk, long gtot) {
  int rc = 0;
  if (ntot > cap_n_ || nbtot > cap_nb_ || ctot > cap_c_ || ninst > cap_inst_ ||
      nblk > cap_blk_ || gtot > cap_g_) {
    double *hq_keep = d_.hq;
    free_all();
    d_.hq = hq_keep;
    cap_n_ = std::max(ntot, 1);
    cap_nb_ = std::max(nbtot, 1);
    cap_c_ = std::max(ctot, 1);
    cap_inst_ = std::max(ninst, 1);
    cap_blk_ = std::max(nblk, 1);
    cap_g_ = std::max(gtot, 1L);
    const size_t N = cap_n_, NB = cap_nb_, C = cap_c_ + 1;
    for (int b = 0; b < 2; ++b) {
      rc |= dalloc(allocs_, &d_.x[b], N);
      rc |= dalloc(allocs_, &d_.y[b], N);
      rc |= dalloc(allocs_, &d_.vx[b], N);
      rc |= dalloc(allocs_, &d_.vy[b], N);
      rc |= dalloc(allocs_, &d_.h[b], N);
      rc |= dalloc(allocs_, &d_.A[b], N);
      rc |= dalloc(allocs_, &d_.rho[b], N);
      rc |= dalloc(allocs_, &d_.hsml[b], N);
    }
    rc |= dalloc(allocs_, &d_.mass, N);
    rc |= dalloc(allocs_, &d_.orig, N);
    rc |= dalloc(allocs_, &d_.inst, N);
    rc |= dalloc(allocs_, &d_.cellid, N);
    rc |= dalloc(allocs_, &d_.slot, N);
    rc |= dalloc(allocs_, &d_.cl, N);
    rc |= dalloc(allocs_, &d_.clpos, N);
    rc |= dalloc(allocs_, &d_.own_cnt, N);
    rc |= dalloc(allocs_, &d_.moff, N);
    rc |= dalloc(allocs_, &d_.xs, N);
    rc |= dalloc(allocs_, &d_.ys, N);
    rc |= dalloc(allocs_, &d_.hs, N);
    rc |= dalloc(allocs_, &d_.vxs, N);
    rc |= dalloc(allocs_, &d_.vys, N);
    rc |= dalloc(allocs_, &d_.ms, N);
    rc |= dalloc(allocs_, &d_.rs, N);
    rc |= dalloc(allocs_, &d_.cells, N);
    rc |= dalloc(allocs_, &d_.pf, N);
    rc |= dalloc(allocs_, &d_.nbr, (size_t)kNbrMax * N);
    rc |= dalloc(allocs_, &d_.nnb, N);
    rc |= dalloc(allocs_, &d_.part, 5 * (size_t)cap_blk_);
    rc |= dalloc(allocs_, &d_.bflag, C);
    rc |= dalloc(allocs_, &d_.count, C);
    rc |= dalloc(allocs_, &d_.cell_start, C);
    rc |= dalloc(allocs_, &d_.pair_count, C);
    rc |= dalloc(allocs_, &d_.pair_start, C);
    rc |= dalloc(allocs_, &d_.cell_intra, C);
    double *sg = nullptr;
    rc |= dalloc(allocs_, &sg, N); d_.sig11 = sg;
    rc |= dalloc(allocs_, &sg, N); d_.sig12 = sg;
    rc |= dalloc(allocs_, &sg, N); d_.sig22 = sg;
    rc |= dalloc(allocs_, &sg, N); d_.div = sg;
    rc |= dalloc(allocs_, &d_.diag, 6 * N);
    rc |= dalloc(allocs_, &d_.sig_out, 3 * N);
    rc |= dalloc(allocs_, &d_.tend, 6 * N);
    rc |= dalloc(allocs_, &d_.xb, NB);
    rc |= dalloc(allocs_, &d_.yb, NB);
    rc |= dalloc(allocs_, &d_.massb, NB);
    rc |= dalloc(allocs_, &d_.bpre, 9 * NB);
    rc |= dalloc(allocs_, &d_.bbase, NB);
    rc |= dalloc(allocs_, &d_.bcount, NB);
    rc |= dalloc(allocs_, &d_.bsorted, NB);
    rc |= dalloc(allocs_, &d_.grid, (size_t)cap_g_);
    rc |= dalloc(allocs_, &d_.meta, (size_t)cap_inst_);
    rc |= dalloc(allocs_, &d_.dyn, (size_t)cap_inst_);
    rc |= dalloc(allocs_, &d_.blk, (size_t)cap_blk_);
    cap_scan_ = cdiv(C + 1, kScanItems) + 1;
    if (hipMalloc((void **)&scan_tmp_, cap_scan_ * sizeof(int)) != hipSuccess) rc = -6;
    if (rc) return -6;
    if (hipMemsetAsync(d_.count, 0, C * sizeof(int), st_) != hipSuccess) return -2;
    if (gexec_) { hipGraphExecDestroy(gexec_); gexec_ = nullptr; }
    if (graph_) { hipGraphDestroy(graph_); graph_ = nullptr; }
  }
  if (qtot > cap_q_) {
    double *nq = nullptr;
    if (hipMalloc((void **)&nq, (size_t)qtot * sizeof(double)) != hipSuccess) return -6;
    if (hipMemsetAsync(nq, 0, (size_t)qtot * sizeof(double), st_) != hipSuccess) return -2;
    if (d_.hq && cap_q_ > 0)
      if (hipMemcpyAsync(nq, d_.hq, (size_t)cap_q_ * sizeof(double), hipMemcpyDeviceToDevice,
                         st_) != hipSuccess)
        return -2;
    if (hipStreamSynchronize(st_) != hipSuccess) return -2;
    if (d_.hq) hipFree(d_.hq);
    d_.hq = nq;
    cap_q_ = qtot;
    if (gexec_) { hipGraphExecDestroy(gexec_); gexec_ = nullptr; }
    if (graph_) { hipGraphDestroy(graph_); graph_ = nullptr; }
  }
  return 0;
}

int Engine::upload(const std::vector<InstanceInput> &ins, bool keep_store) {
  const int ninst = (int)ins.size();
  if (ninst < 1) return -1;
  std::vector<InstMeta> meta(ninst);
  int ntot = 0, nbtot = 0, ctot = 0, qtot = 0, nblk = 0;
  long gtot = 0;
  any_boundary_ = false;
  for (int k = 0; k < ninst; ++k) {
    const InstanceInput &in = ins[k];
    if (in.n < 1 || in.nb < 0 || in.nb > kMaxBoundaryPerInst) return -1;
    if (!in.x || !in.vx || !in.mass || !in.rho || !in.hsml || !in.hth || !in.A) return -1;
    if (in.nb > 0 && (!in.xb || !in.massb)) return -1;
    InstMeta &m = meta[k];
    std::memset(&m, 0, sizeof(m));
    m.p_off = ntot; m.n = in.n;
    m.blk_off = nblk; m.nblk = cdiv(in.n, kTPB);
    m.b_off = nbtot; m.nb = in.nb;
    long cells = in.c_cap;
    if (cells <= 0) {
      long g = grid_cells(in.n, in.x, in.hsml, in.nb, in.xb, 1.0);
      if (g < 0) return -4;
      // room for smoothing lengths shrinking to 70 % and a drifting domain
      long g2 = grid_cells(in.n, in.x, in.hsml, in.nb, in.xb, 0.7);
      cells = std::max(g2 * 2, g + 1024);
    }
    if (cells > (1L << 28)) return -3;
    m.c_off = ctot; m.c_cap = (int)cells;
    int q = in.q_cap > 0 ? in.q_cap : (in.nb > 0 ? std::max(64 * in.nb, 1024) : 0);
    m.q_off = qtot; m.q_cap = q;
    m.tse = in.tse;
    m.wind_mode = in.wind_mode;
    m.wind_u = in.wind[0]; m.wind_v = in.wind[1];
    m.cur_u = in.current[0]; m.cur_v = in.current[1];
    m.max_hsml = in.max_hsml;
    if (in.wind_mode == 1) {
      if (in.g_nx < 2 || in.g_ny < 2 || !in.g_u || !in.g_v) return -1;
      m.g_off = (int)gtot; m.g_nx = in.g_nx; m.g_ny = in.g_ny;
      m.g_x0 = in.g_x0; m.g_y0 = in.g_y0; m.g_dx = in.g_dx; m.g_dy = in.g_dy;
      gtot += 2L * in.g_nx * in.g_ny;
    }
    m.bxmin = m.bymin = 1e300;
    m.bxmax = m.bymax = -1e300;
    for (int b = 0; b < in.nb; ++b) {
      m.bxmin = std::min(m.bxmin, in.xb[2 * b]); m.bxmax = std::max(m.bxmax, in.xb[2 * b]);
      m.bymin = std::min(m.bymin, in.xb[2 * b + 1]); m.bymax = std::max(m.bymax, in.xb[2 * b + 1]);
    }
    if (in.nb > 0) any_boundary_ = true;
    ntot += in.n; nbtot += in.nb; ctot += m.c_cap; qtot += q;
    nblk += cdiv(in.n, kTPB);
  }
  bool layout_same = keep_store && meta_.size() == meta.size();
  if (layout_same)
    for (int k = 0; k < ninst; ++k)
      if (meta_[k].q_off != meta[k].q_off || meta_[k].q_cap > meta[k].q_cap) layout_same = false;
  if (layout_same) {
    qtot = 0;
    for (int k = 0; k < ninst; ++k) {
      meta[k].q_cap = std::max(meta[k].q_cap, meta_[k].q_cap);
      meta[k].q_off = qtot;
      qtot += meta[k].q_cap;
    }
  }
  int rc = ensure(ntot, nbtot, ctot, qtot, ninst, nblk, gtot);
  if (rc) return rc;
  if (!layout_same && d_.hq)
    HIPCHK(hipMemsetAsync(d_.hq, 0, (size_t)cap_q_ * sizeof(double), st_));

  // host staging in SoA
  std::vector<double> hx(ntot), hy(ntot), hvx(ntot), hvy(ntot), hm(ntot), hr(ntot), hs(ntot),
      hh(ntot), hA(ntot);
  std::vector<int> horig(ntot), hinst(ntot);
  std::vector<double> hxb(std::max(nbtot, 1)), hyb(std::max(nbtot, 1)), hmb(std::max(nbtot, 1));
  std::vector<double> hgrid(std::max(gtot, 1L));
  std::vector<BlockEntry> hblk;
  std::vector<InstDyn> hdyn(ninst);
  for (int k = 0; k < ninst; ++k) {
    const InstanceInput &in = ins[k];
    const InstMeta &m = meta[k];
    for (int i = 0; i < in.n; ++i) {
      int p = m.p_off + i;
      hx[p] = in.x[2 * i]; hy[p] = in.x[2 * i + 1];
      hvx[p] = in.vx[2 * i]; hvy[p] = in.vx[2 * i + 1];
      hm[p] = in.mass[i]; hr[p] = in.rho[i]; hs[p] = in.hsml[i];
      hh[p] = in.hth[i]; hA[p] = in.A[i];
      horig[p] = i; hinst[p] = k;
    }
    for (int b = 0; b < in.nb; ++b) {
      hxb[m.b_off + b] = in.xb[2 * b];
      hyb[m.b_off + b] = in.xb[2 * b + 1];
      hmb[m.b_off + b] = in.massb[b];
    }
    if (in.wind_mode == 1) {
      long nn = (long)in.g_nx * in.g_ny;
      std::memcpy(&hgrid[m.g_off], in.g_u, nn * sizeof(double));
      std::memcpy(&hgrid[m.g_off + nn], in.g_v, nn * sizeof(double));
    }
    for (int f = 0; f < in.n; f += kTPB) hblk.push_back(BlockEntry{k, f, (int)hblk.size(), 0});
    InstDyn &y = hdyn[k];
    std::memset(&y, 0, sizeof(y));
    y.dt = in.dt; y.time = in.time; y.itimestep = in.itimestep;
  }
  auto up = [&](void *dst, const void *src, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st_);
  };
  const size_t N8 = (size_t)ntot * 8;
  HIPCHK(up(d_.x[0], hx.data(), N8));
  HIPCHK(up(d_.y[0], hy.data(), N8));
  HIPCHK(up(d_.vx[0], hvx.data(), N8));
  HIPCHK(up(d_.vy[0], hvy.data(), N8));
  HIPCHK(up(d_.mass, hm.data(), N8));
  HIPCHK(up(d_.rho[0], hr.data(), N8));
  HIPCHK(up(d_.hsml[0], hs.data(), N8));
  HIPCHK(up(d_.h[0], hh.data(), N8));
  HIPCHK(up(d_.A[0], hA.data(), N8));
  HIPCHK(up(d_.orig, horig.data(), (size_t)ntot * 4));
  HIPCHK(up(d_.inst, hinst.data(), (size_t)ntot * 4));
  if (nbtot > 0) {
    HIPCHK(up(d_.xb, hxb.data(), (size_t)nbtot * 8));
    HIPCHK(up(d_.yb, hyb.data(), (size_t)nbtot * 8));
    HIPCHK(up(d_.massb, hmb.data(), (size_t)nbtot * 8));
  }
  if (gtot > 0) HIPCHK(up(d_.grid, hgrid.data(), (size_t)gtot * 8));
  HIPCHK(up(d_.meta, meta.data(), sizeof(InstMeta) * ninst));
  HIPCHK(up(d_.dyn, hdyn.data(), sizeof(InstDyn) * ninst));
  // Workgroups are dispatched to the 8 XCDs round-robin by block id; the
  // physical table gives each XCD a contiguous run of logical blocks.
  {
    const int nb_ = (int)hblk.size(), q = nb_ / 8, r = nb_ % 8;
    std::vector<BlockEntry> phys(nb_);
    for (int b = 0; b < nb_; ++b) {
      const int x = b % 8;
      phys[b] = hblk[x * q + std::min(x, r) + b / 8];
    }
    HIPCHK(up(d_.blk, phys.data(), sizeof(BlockEntry) * nb_));
  }
  HIPCHK(hipMemsetAsync(d_.bflag, 0xff, sizeof(int) * ((size_t)ctot + 1), st_));
  HIPCHK(hipStreamSynchronize(st_));
  d_.ntot = ntot; d_.nbtot = nbtot; d_.ctot = ctot; d_.qtot = qtot; d_.ninst = ninst;
  d_.nblk = (int)hblk.size();
  meta_ = meta;
  if (gexec_) { hipGraphExecDestroy(gexec_); gexec_ = nullptr; }
  if (graph_) { hipGraphDestroy(graph_); graph_ = nullptr; }
  // grid of the first evaluation
  launch_partial(d_, 0, st_);
  launch_finalize(d_, 0, st_);

  // size the boundary-pair stores from a probe of the first evaluation
  if (any_boundary_) {
    std::vector<int> bnip;
    rc = probe_bnip(bnip);
    if (rc) return rc;
    bool grow = false;
    for (int k = 0; k < ninst; ++k)
      if (bnip[k] > meta_[k].q_cap) grow = true;
    if (grow) {
      int q = 0;
      std::vector<double> old_hq;
      for (int k = 0; k < ninst; ++k) {
        int cap = std::max(meta_[k].q_cap, bnip[k] > meta_[k].q_cap ? 2 * bnip[k] + 1024 : 0);
        meta_[k].q_off = q;
        meta_[k].q_cap = cap;
        q += cap;
      }
      if (ninst == 1) {
        rc = ensure(ntot, nbtot, ctot, q, ninst, nblk, gtot);  // keeps old entries
      } else {
        hipFree(d_.hq);
        d_.hq = nullptr;
        cap_q_ = 0;
        rc = ensure(ntot, nbtot, ctot, q, ninst, nblk, gtot);
      }
      if (rc) return rc;
      d_.qtot = q;
      HIPCHK(up(d_.meta, meta_.data(), sizeof(InstMeta) * ninst));
    }
    // the probe may have flagged the store capacity: restart clean
    HIPCHK(up(d_.dyn, hdyn.data(), sizeof(InstDyn) * ninst));
    launch_partial(d_, 0, st_);
    launch_finalize(d_, 0, st_);
    HIPCHK(hipStreamSynchronize(st_));
  }
  return 0;
}

template <class F>
void Engine::L(const char *name, F &&f) {
  if (!prof) {
    f();
    return;
  }
  Profiler::Rec r;
  r.name = name;
  hipEventCreate(&r.a);
  hipEventCreate(&r.b);
  hipEventRecord(r.a, st_);
  f();
  hipEventRecord(r.b, st_);
  prof->recs.push_back(r);
}

void Engine::enqueue_prefix(int buf, bool fork_boundary) {
  const DevState &d = d_;
  L("count", [&] { launch_count(d, buf, st_); });
  L("scan", [&] { launch_scan(d.count, d.ctot, scan_tmp_, d.cell_start, 1, st_); });
  L("scatter", [&] { launch_scatter(d, st_); });
  L("cellsort", [&] { launch_cellsort(d, buf, st_); });
  if (!any_boundary_) return;
  if (fork_boundary && !prof) {
    // the boundary prep only feeds materialize and pass B: it runs on a
    // second stream (a parallel branch of the captured graph) beside pass A
    hipEventRecord(ev_fork_, st_);
    hipStreamWaitEvent(side_, ev_fork_, 0);
    launch_bprep(d, side_);
    hipEventRecord(ev_join_, side_);
  } else {
    L("bprep", [&] { launch_bprep(d, st_); });
  }
}

void Engine::enqueue_eval(int buf, int phase, bool want_diag) {
  const DevState &d = d_;
  const bool fork = any_boundary_ && !prof;
  enqueue_prefix(buf, fork);
  L("passA", [&] { launch_passA(d, buf, precision_, want_diag ? 1 : 0, st_); });
  if (any_boundary_ || phase == 2) {
    // pair ranks: needed by the boundary kernel's hsml_ij read, and for nip
    L("pairtot", [&] { launch_pairtot(d, st_); });
    L("scan", [&] { launch_scan(d.pair_count, d.ctot, scan_tmp_, d.pair_start, 0, st_); });
  }
  if (fork) hipStreamWaitEvent(st_, ev_join_, 0);
  if (any_boundary_) L("materialize", [&] { launch_materialize(d, st_); });
  L("passB", [&] { launch_passB(d, precision_, phase, st_); });
}

void Engine::enqueue_step() {
  enqueue_eval(0, 0, false);
  L("finalize", [&] { launch_finalize(d_, 0, st_); });
  enqueue_eval(1, 1, diag_);
  L("finalize", [&] { launch_finalize(d_, 1, st_); });
}

int Engine::probe_bnip(std::vector<int> &bnip) {
  enqueue_prefix(0, false);
  HIPCHK(hipStreamSynchronize(st_));
  std::vector<InstDyn> hd(d_.ninst);
  HIPCHK(hipMemcpy(hd.data(), d_.dyn, sizeof(InstDyn) * d_.ninst, hipMemcpyDeviceToHost));
  bnip.resize(d_.ninst);
  for (int k = 0; k < d_.ninst; ++k) {
    if (hd[k].err & ~kErrCapacity) return (hd[k].err & kErrNan) ? -4 : -3;
    bnip[k] = hd[k].bnip;
  }
  return 0;
}

int Engine::eval_single(bool want_diag) {
  enqueue_eval(0, 2, want_diag);
  return sync_and_check();
}

int Engine::graph_build() {
  if (gexec_) return 0;
  HIPCHK(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
  enqueue_step();
  HIPCHK(hipStreamEndCapture(st_, &graph_));
  HIPCHK(hipGraphInstantiate(&gexec_, graph_, nullptr, nullptr, 0));
  return 0;
}

int Engine::steps(int nsteps, bool use_graph) {
  if (nsteps < 0) return -1;
  if (use_graph && !prof) {
    int rc = graph_build();
    if (rc) return rc;
    for (int s = 0; s < nsteps; ++s) HIPCHK(hipGraphLaunch(gexec_, st_));
  } else {
    for (int s = 0; s < nsteps; ++s) enqueue_step();
  }
  return 0;
}

int Engine::sync_and_check() {
  HIPCHK(hipStreamSynchronize(st_));
  std::vector<InstDyn> hd(d_.ninst);
  HIPCHK(hipMemcpy(hd.data(), d_.dyn, sizeof(InstDyn) * d_.ninst, hipMemcpyDeviceToHost));
  int worst = 0;
  for (auto &y : hd) {
    if (y.err & kErrNan) return -4;
    if (y.err) worst = -3;
  }
  return worst;
}

int Engine::download_scalars(int k, double *dt, double *time, int *itimestep, int *nip,
                             int *bnip, int *err) {
  if (k < 0 || k >= d_.ninst) return -1;
  InstDyn y;
  HIPCHK(hipMemcpy(&y, d_.dyn + k, sizeof(y), hipMemcpyDeviceToHost));
  if (dt) *dt = y.dt;
  if (time) *time = y.time;
  if (itimestep) *itimestep = y.itimestep;
  if (bnip) *bnip = y.bnip;
  if (err) *err = y.err;
  if (nip) {
    const InstMeta &m = meta_[k];
    int a = 0, b = 0;
    HIPCHK(hipMemcpy(&a, d_.pair_start + m.c_off, 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&b, d_.pair_start + m.c_off + m.c_cap, 4, hipMemcpyDeviceToHost));
    *nip = b - a;
  }
  return 0;
}

static void interleave(const std::vector<double> &a, const std::vector<double> &b, double *out,
                       int n) {
  for (int i = 0; i < n; ++i) {
    out[2 * i] = a[i];
    out[2 * i + 1] = b[i];
  }
}

int Engine::download_state(int k, double *x, double *vx, double *rho, double *hsml, double *hth,
                           double *A) {
  if (k < 0 || k >= d_.ninst) return -1;
  HIPCHK(hipStreamSynchronize(st_));
  const InstMeta &m = meta_[k];
  const int n = m.n;
  std::vector<double> a(n), b(n);
  auto get = [&](double *dst, const double *src) {
    return hipMemcpy(dst, src + m.p_off, (size_t)n * 8, hipMemcpyDeviceToHost);
  };
  if (x) {
    HIPCHK(get(a.data(), d_.x[0]));
    HIPCHK(get(b.data(), d_.y[0]));
    interleave(a, b, x, n);
  }
  if (vx) {
    HIPCHK(get(a.data(), d_.vx[0]));
    HIPCHK(get(b.data(), d_.vy[0]));
    interleave(a, b, vx, n);
  }
  if (rho) HIPCHK(get(rho, d_.rho[0]));
  if (hsml) HIPCHK(get(hsml, d_.hsml[0]));
  if (hth) HIPCHK(get(hth, d_.h[0]));
  if (A) HIPCHK(get(A, d_.A[0]));
  return 0;
}

int Engine::download_tend(int k, double *dxdt, double *dvxdt, double *dhdt, double *dAdt) {
  if (k < 0 || k >= d_.ninst) return -1;
  const InstMeta &m = meta_[k];
  const int n = m.n, N = d_.ntot;
  std::vector<double> a(n), b(n);
  auto get = [&](double *dst, int comp) {
    return hipMemcpy(dst, d_.tend + (size_t)comp * N + m.p_off, (size_t)n * 8,
                     hipMemcpyDeviceToHost);
  };
  if (dxdt) {
    HIPCHK(get(a.data(), 0));
    HIPCHK(get(b.data(), 1));
    interleave(a, b, dxdt, n);
  }
  if (dvxdt) {
    HIPCHK(get(a.data(), 2));
    HIPCHK(get(b.data(), 3));
    interleave(a, b, dvxdt, n);
  }
  if (dhdt) HIPCHK(get(dhdt, 4));
  if (dAdt) HIPCHK(get(dAdt, 5));
  return 0;
}

int Engine::download_diag(int k, double *eps3, double *zeta, double *eta, double *p,
                          double *sig3) {
  if (k < 0 || k >= d_.ninst) return -1;
  HIPCHK(hipStreamSynchronize(st_));
  const InstMeta &m = meta_[k];
  const int n = m.n, N = d_.ntot;
  std::vector<double> c0(n), c1(n), c2(n);
  auto get = [&](double *dst, int comp) {
    return hipMemcpy(dst, d_.diag + (size_t)comp * N + m.p_off, (size_t)n * 8,
                     hipMemcpyDeviceToHost);
  };
  if (eps3) {
    HIPCHK(get(c0.data(), 0));
    HIPCHK(get(c1.data(), 1));
    HIPCHK(get(c2.data(), 2));
    for (int i = 0; i < n; ++i) {
      eps3[3 * i] = c0[i]; eps3[3 * i + 1] = c1[i]; eps3[3 * i + 2] = c2[i];
    }
  }
  if (zeta) HIPCHK(get(zeta, 3));
  if (eta) HIPCHK(get(eta, 4));
  if (p) HIPCHK(get(p, 5));
  if (sig3) {
    HIPCHK(hipMemcpy(c0.data(), d_.sig_out + m.p_off, n * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(c1.data(), d_.sig_out + (size_t)N + m.p_off, n * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(c2.data(), d_.sig_out + 2 * (size_t)N + m.p_off, n * 8,
                     hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) {
      sig3[3 * i] = c0[i]; sig3[3 * i + 1] = c1[i]; sig3[3 * i + 2] = c2[i];
    }
  }
  return 0;
}

int Engine::export_pairs(int cap, int *nip, int *pi, int *pj, double *rij, double *hij,
                         int bcap, int *bnip, int *bpi, int *bpj, double *brij) {
  // requires a prefix + passA (own counts) + pair ranks for state S
  const DevState &d = d_;
  enqueue_prefix(0, false);
  launch_passA(d, 0, precision_, 0, st_);
  launch_pairtot(d, st_);
  launch_scan(d.pair_count, d.ctot, scan_tmp_, d.pair_start, 0, st_);
  int rc = sync_and_check();
  if (rc) return rc;
  int n_ip = 0, n_bp = 0, err = 0;
  rc = download_scalars(0, nullptr, nullptr, nullptr, &n_ip, &n_bp, &err);
  if (rc) return rc;
  if (!any_boundary_) n_bp = 0;
  *nip = n_ip;
  *bnip = n_bp;
  int *dpi = nullptr, *dpj = nullptr, *dbi = nullptr, *dbj = nullptr;
  double *dr = nullptr, *dh = nullptr, *dbr = nullptr;
  const int c1 = std::max(std::min(cap, n_ip), 1), c2 = std::max(std::min(bcap, n_bp), 1);
  bool ok = hipMalloc((void **)&dpi, c1 * 4) == hipSuccess &&
            hipMalloc((void **)&dpj, c1 * 4) == hipSuccess &&
            hipMalloc((void **)&dr, c1 * 8) == hipSuccess &&
            hipMalloc((void **)&dh, c1 * 8) == hipSuccess &&
            hipMalloc((void **)&dbi, c2 * 4) == hipSuccess &&
            hipMalloc((void **)&dbj, c2 * 4) == hipSuccess &&
            hipMalloc((void **)&dbr, c2 * 8) == hipSuccess;
  if (ok) {
    launch_export_pairs(d, 0, std::min(cap, n_ip), dpi, dpj, dr, dh, st_);
    launch_export_bpairs(d, 0, any_boundary_ ? meta_[0].nb : 0, std::min(bcap, n_bp), dbi, dbj,
                         dbr, st_);
    ok = hipStreamSynchronize(st_) == hipSuccess;
    int k1 = std::min(cap, n_ip), k2 = std::min(bcap, n_bp);
    if (ok && k1 > 0) {
      if (pi) ok &= hipMemcpy(pi, dpi, k1 * 4, hipMemcpyDeviceToHost) == hipSuccess;
      if (pj) ok &= hipMemcpy(pj, dpj, k1 * 4, hipMemcpyDeviceToHost) == hipSuccess;
      if (rij) ok &= hipMemcpy(rij, dr, k1 * 8, hipMemcpyDeviceToHost) == hipSuccess;
      if (hij) ok &= hipMemcpy(hij, dh, k1 * 8, hipMemcpyDeviceToHost) == hipSuccess;
    }
    if (ok && k2 > 0) {
      if (bpi) ok &= hipMemcpy(bpi, dbi, k2 * 4, hipMemcpyDeviceToHost) == hipSuccess;
      if (bpj) ok &= hipMemcpy(bpj, dbj, k2 * 4, hipMemcpyDeviceToHost) == hipSuccess;
      if (brij) ok &= hipMemcpy(brij, dbr, k2 * 8, hipMemcpyDeviceToHost) == hipSuccess;
    }
  }
  hipFree(dpi); hipFree(dpj); hipFree(dr); hipFree(dh);
  hipFree(dbi); hipFree(dbj); hipFree(dbr);
  if (!ok) return -2;
  if (n_ip > cap || n_bp > bcap) return -3;
  return 0;
}

}  // namespace simp
