// simp_engine.hip -- packed multi-instance engine: device buffers, uploads,
// the per-evaluation kernel pipeline, the captured step graph, downloads.
#include "simp_engine.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <unordered_map>

#include "simp_kernels.h"

namespace simp {

#define HIPCHK(x)                       \
  do {                                  \
    if ((x) != hipSuccess) return -2;   \
  } while (0)

static inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// an instance with a switch value beyond the reference's defaults: the
// kernels then run their variants that read each instance's values
static inline bool nondefault(const InstMeta &m) {
  return m.den_alg != 3 || m.pa_sph == 2 || m.eos_alg != 3 || m.rheology != 1 || m.skf != 6 ||
         m.tse == 1 || m.sle != 2 || m.xsph || m.avis || m.bfa != 1 || m.fric || !m.bpres ||
         !m.wind_on || !m.water_on || m.novisc;
}

// The switch fields of an instance's record.
static void meta_switches(InstMeta &m, const InstanceInput &in) {
  m.den_alg = in.den_alg;
  m.pa_sph = in.pa_sph;
  // without the pressure force eos is not called and p stays the ice
  // strength: the Hibler pressure (internal_stress.f90:21-39), or, with the
  // viscosity off too, initialize's zero (eos_alg 0)
  m.eos_alg = in.pressure_force ? in.eos_alg : (in.viscosity ? 1 : 0);
  m.rheology = in.rheology;
  m.skf = in.skf;
  m.sle = in.sle;
  m.fric = in.boundary_friction ? 1 : 0;
  m.bpres = in.boundary_pressure ? 1 : 0;
  m.wind_on = in.wind_force ? 1 : 0;
  m.water_on = in.water_force ? 1 : 0;
  m.novisc = in.viscosity ? 0 : 1;
  m.nostrain = (!in.viscosity && !in.pressure_force) ? 1 : 0;
  m.xsph = in.xsph ? 1 : 0;
  // regularization runs for XSPH or the artificial viscosity, and then
  // always with the viscosity art_visc_alg names (regularization.f90:29-35)
  m.avis = (in.xsph || in.art_visc) ? in.art_visc_alg : 0;
  // c_yang is computed only for these switches (sound_speed.f90:21)
  m.needc = (in.art_visc || in.tse == 1 || in.boundary_friction) ? 1 : 0;
  m.avdt = (in.art_visc && in.tse == 1) ? 1 : 0;
  m.bfa = in.bfa;
  m.endp = (in.sle == 1 || m.avdt) ? 1 : 0;
}

// device bytes of the arrays the regularization and smoothing-length
// switches add, per particle of an instance that uses them
static size_t ext_bytes_per_particle(const InstanceInput &in) {
  size_t b = 0;
  if (in.sle != 2) b += 8 + 2 * 8 + 2 * 8;                   // drp, masst, hnew / cfl
  if (in.xsph || in.art_visc) b += 8 + kNbrMax * 4 + (kNbrMax / 2) * 12;   // lv, pidx, pairs
  return b;
}

// ------------------------------------------------------------ host sizing

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

static long cell_capacity(const InstanceInput &in) {
  long g = grid_cells(in.n, in.x, in.hsml, in.nb, in.xb, 1.0);
  if (g < 0) return -1;
  // room for smoothing lengths shrinking to 70 % and a drifting domain
  long g2 = grid_cells(in.n, in.x, in.hsml, in.nb, in.xb, 0.7);
  return std::max(g2 * 2, g + 1024);
}

// Number of boundary pairs of the first evaluation (the boundary part of
// bucket_search, nnps.f90:451-493), counted on the host to size the
// hsml_ij store.
static long boundary_pairs(const InstanceInput &in) {
  if (in.nb == 0) return 0;
  double hmin = in.hsml[0];
  for (int i = 1; i < in.n; ++i) hmin = std::min(hmin, in.hsml[i]);
  const double bin = 2 * hmin;
  std::unordered_map<long long, std::vector<int>> cells;
  auto key = [](long cx, long cy) { return (long long)cx * 2147483647LL + cy; };
  for (int i = 0; i < in.n; ++i)
    cells[key(std::lround(in.x[2 * i] / bin), std::lround(in.x[2 * i + 1] / bin))].push_back(i);
  long total = 0;
  for (int b = 0; b < in.nb; ++b) {
    const double bx = in.xb[2 * b], by = in.xb[2 * b + 1];
    const long cx = std::lround(bx / bin), cy = std::lround(by / bin);
    for (long yy = cy - 1; yy <= cy + 1; ++yy)
      for (long xx = cx - 1; xx <= cx + 1; ++xx) {
        auto it = cells.find(key(xx, yy));
        if (it == cells.end()) continue;
        for (int i : it->second) {
          const double dx = bx - in.x[2 * i], dy = by - in.x[2 * i + 1];
          if (std::sqrt(dx * dx + dy * dy) <= kBLength) ++total;
        }
      }
  }
  return total;
}

static int store_capacity(const InstanceInput &in) {
  if (in.nb == 0) return 0;
  const long bp = boundary_pairs(in);
  return (int)std::min<long>(2 * bp + 1024, 1L << 30);
}

// bytes per particle / boundary particle / cell of the device layout
constexpr size_t kBytesPerParticle =
    4 * 8 * 8 +             // x .. hsml, two sets of S and T
    2 * (8 + 4 + 4) +       // mass, orig, cellid per set
    2 * 4 + 2 * 4 + 8 + 4 + // slot, cl; own_cnt, moff; gpre, osig
    16 +                    // pf
    (kNbrMax + 1) * 4 + 4 + // nbr (a spare row), nnb
    kNbrMax * 4 + 4 +       // vnb, vnc
    5 * 8 + 3 * 8 +         // sig11 (x2: srec)/12/22, div; kz, dnum, dsum
    6 * 8 + 3 * 8 + 6 * 8;  // diag, sig_out, tend
constexpr size_t kBytesPerBoundary = 3 * 8 + 9 * 4 + 2 * 4 + 9 * 4 + 9 * 4 + 4;
constexpr size_t kBytesPerCell = 9 * 4 + 8 + 8;

size_t instance_footprint(const InstanceInput &in) {
  long c = cell_capacity(in);
  if (c < 0) c = 0;
  return (kBytesPerParticle + ext_bytes_per_particle(in)) * (size_t)in.n +
         (kBytesPerBoundary + (in.bfa == 2 ? 16 : 0)) * (size_t)in.nb +
         kBytesPerCell * (size_t)c + 16 * (size_t)store_capacity(in) +
         16 * (size_t)in.g_nx * (size_t)in.g_ny + 4096;
}

// ------------------------------------------------------------ allocation

// An allocator is called as alloc(&ptr, bytes) and returns nonzero when it
// cannot provide the bytes.
template <class T, class A>
static int dalloc(A &alloc, T **p, size_t count) {
  return alloc((void **)p, std::max<size_t>(count, 1) * sizeof(T));
}

static void free_list(std::vector<void *> &list) {
  for (void *p : list) hipFree(p);
  list.clear();
}

// Every buffer of a layout with the given totals (the hsml_ij store, its
// pair slots and kernel_zero apart).  Returns nonzero when the allocator
// fails.
template <class A>
static int alloc_state(DevState &d, A &&alloc, int **scan_tmp, size_t N, size_t NB, size_t C,
                       size_t BLK, size_t INST, size_t G) {
  int rc = 0;
  N = std::max<size_t>(N, 1);
  NB = std::max<size_t>(NB, 1);
  C = std::max<size_t>(C, 1) + 1;
  for (int b = 0; b < 4; ++b) {
    rc |= dalloc(alloc, &d.x[b], N);
    rc |= dalloc(alloc, &d.y[b], N);
    rc |= dalloc(alloc, &d.vx[b], N);
    rc |= dalloc(alloc, &d.vy[b], N);
    rc |= dalloc(alloc, &d.h[b], N);
    rc |= dalloc(alloc, &d.A[b], N);
    rc |= dalloc(alloc, &d.rho[b], N);
    rc |= dalloc(alloc, &d.hsml[b], N);
  }
  for (int b = 0; b < 2; ++b) {
    rc |= dalloc(alloc, &d.mass[b], N);
    rc |= dalloc(alloc, &d.orig[b], N);
    rc |= dalloc(alloc, &d.cellid[b], N);
  }
  rc |= dalloc(alloc, &d.slot, N);
  rc |= dalloc(alloc, &d.cl, N);
  rc |= dalloc(alloc, &d.own_cnt, N);
  rc |= dalloc(alloc, &d.moff, N);
  rc |= dalloc(alloc, &d.gpre, N);
  rc |= dalloc(alloc, &d.osig, N);
  rc |= dalloc(alloc, &d.pf, N);
  rc |= dalloc(alloc, &d.nbr, (size_t)(kNbrMax + 1) * N);   // row kNbrMax: spare
  rc |= dalloc(alloc, &d.nnb, N);
  rc |= dalloc(alloc, &d.vnb, (size_t)kNbrMax * N);
  rc |= dalloc(alloc, &d.vnc, N);

  rc |= dalloc(alloc, &d.part, kParts * BLK);
  rc |= dalloc(alloc, &d.bflag, C);
  rc |= dalloc(alloc, &d.count, C);
  rc |= dalloc(alloc, &d.fill, C);
  rc |= dalloc(alloc, &d.cinst, C);
  rc |= dalloc(alloc, &d.cell_start, C);
  rc |= dalloc(alloc, &d.pair_count, C);
  rc |= dalloc(alloc, &d.pair_start, C);
  rc |= dalloc(alloc, &d.cell_intra, C);
  rc |= dalloc(alloc, &d.cellh, C);
  rc |= dalloc(alloc, &d.rebuild, 1);
  rc |= dalloc(alloc, &d.bar, kBarWords);
  double *sg = nullptr;
  rc |= dalloc(alloc, &sg, 2 * N); d.sig11 = sg;
  d.srec = (float4 *)sg;
  rc |= dalloc(alloc, &sg, N); d.sig12 = sg;
  rc |= dalloc(alloc, &sg, N); d.sig22 = sg;
  rc |= dalloc(alloc, &sg, N); d.div = sg;
  rc |= dalloc(alloc, &d.dnum, N);
  rc |= dalloc(alloc, &d.dsum, N);
  rc |= dalloc(alloc, &d.diag, 6 * N);
  rc |= dalloc(alloc, &d.sig_out, 3 * N);
  rc |= dalloc(alloc, &d.tend, 6 * N);
  rc |= dalloc(alloc, &d.xb, NB);
  rc |= dalloc(alloc, &d.yb, NB);
  rc |= dalloc(alloc, &d.massb, NB);
  rc |= dalloc(alloc, &d.bpre, 9 * NB);
  rc |= dalloc(alloc, &d.bbase, NB);
  rc |= dalloc(alloc, &d.bcount, NB);
  rc |= dalloc(alloc, &d.bcnt, 9 * NB);
  rc |= dalloc(alloc, &d.bkey, NB);
  rc |= dalloc(alloc, &d.brun, C);
  rc |= dalloc(alloc, &d.bcellc, C);
  rc |= dalloc(alloc, &d.blist, 9 * NB);
  rc |= dalloc(alloc, &d.grid, std::max<size_t>(G, 1));
  rc |= dalloc(alloc, &d.meta, std::max<size_t>(INST, 1));
  rc |= dalloc(alloc, &d.dyn, std::max<size_t>(INST, 1));
  rc |= dalloc(alloc, &d.blk, std::max<size_t>(BLK, 1));
  rc |= dalloc(alloc, scan_tmp, (size_t)cdiv((long)C + 1, kScanItems) + 1);
  return rc;
}

Engine::Engine(int precision) : precision_(precision) {
  hipStreamCreate(&st_);
  // room for one instance's records, so that laying one out in caller
  // memory allocates nothing on the host either
  meta_.reserve(1);
  ids_.reserve(1);
}

Engine::~Engine() {
  graph_drop();
  if (st_) hipStreamSynchronize(st_);
  if (stage_) hipHostFree(stage_);
  free_list(allocs_);
  if (d_.hq && !arena_) hipFree(d_.hq);
  if (st_) hipStreamDestroy(st_);
}

void Engine::graph_drop() {
  if (gexec_) hipGraphExecDestroy(gexec_);
  if (graph_) hipGraphDestroy(graph_);
  gexec_ = nullptr;
  graph_ = nullptr;
}

int Engine::index_of(int id) const {
  for (size_t k = 0; k < ids_.size(); ++k)
    if (ids_[k] == id) return (int)k;
  return -1;
}

// ----------------------------------------------------------------- repack

int Engine::upload(const std::vector<InstanceInput> &ins, bool keep_store) {
  const int carry = (keep_store && ins.size() == 1 && !meta_.empty()) ? 0 : -1;
  return repack({}, ins, carry, nullptr);
}

int Engine::add(const std::vector<InstanceInput> &ins, std::vector<int> *ids) {
  std::vector<int> keep(meta_.size());
  for (size_t k = 0; k < keep.size(); ++k) keep[k] = (int)k;
  return repack(keep, ins, -1, ids);
}

int Engine::remove(const std::vector<int> &ids) {
  std::vector<int> keep;
  for (size_t k = 0; k < meta_.size(); ++k)
    if (std::find(ids.begin(), ids.end(), ids_[k]) == ids.end()) keep.push_back((int)k);
  return repack(keep, {}, -1, nullptr);
}

int Engine::update(const std::vector<int> &remove_ids, const std::vector<InstanceInput> &ins,
                   std::vector<int> *ids) {
  std::vector<int> keep;
  for (size_t k = 0; k < meta_.size(); ++k)
    if (std::find(remove_ids.begin(), remove_ids.end(), ids_[k]) == remove_ids.end())
      keep.push_back((int)k);
  return repack(keep, ins, -1, ids);
}

// Rebuild the packed layout from the kept instances (old packed indices,
// their live state copied device to device) followed by the added ones
// (copied from the host).  carry_store_from: old index whose hsml_ij store
// and kernel_zero seed the first added instance's.
int Engine::repack(const std::vector<int> &keep, const std::vector<InstanceInput> &add,
                   int carry_store_from, std::vector<int> *new_ids) {
  const int nk = (int)keep.size(), na = (int)add.size(), ninst = nk + na;
  std::vector<InstMeta> nm(ninst);
  std::vector<int> nids(ninst);
  for (int j = 0; j < na; ++j) {
    const InstanceInput &in = add[j];
    if (in.n < 1 || in.nb < 0 || in.nb > kMaxBoundaryPerInst) return -1;
    if (!in.x || !in.vx || !in.mass || !in.rho || !in.hsml || !in.hth || !in.A) return -1;
    if (in.nb > 0 && (!in.xb || !in.massb)) return -1;
    if (in.wind_mode == 1 && (in.g_nx < 2 || in.g_ny < 2 || !in.g_u || !in.g_v)) return -1;
  }
  // sizes and forcing
  for (int j = 0; j < nk; ++j) {
    if (keep[j] < 0 || keep[j] >= (int)meta_.size()) return -1;
    nm[j] = meta_[keep[j]];
    nids[j] = ids_[keep[j]];
  }
  for (int j = 0; j < na; ++j) {
    const InstanceInput &in = add[j];
    InstMeta &m = nm[nk + j];
    std::memset(&m, 0, sizeof(m));
    m.n = in.n;
    m.nb = in.nb;
    const long cells = cell_capacity(in);
    if (cells < 0) return -4;
    if (cells > (1L << 28)) return -3;
    m.c_cap = (int)cells;
    m.q_cap = store_capacity(in);
    if (j == 0 && carry_store_from >= 0 && carry_store_from < (int)meta_.size())
      m.q_cap = std::max(m.q_cap, meta_[carry_store_from].q_cap);
    m.tse = in.tse;
    m.wind_mode = in.wind_mode;
    m.wind_u = in.wind[0]; m.wind_v = in.wind[1];
    m.cur_u = in.current[0]; m.cur_v = in.current[1];
    m.max_hsml = in.max_hsml;
    meta_switches(m, in);
    if (in.wind_mode == 1) {
      m.g_nx = in.g_nx; m.g_ny = in.g_ny;
      m.g_x0 = in.g_x0; m.g_y0 = in.g_y0; m.g_dx = in.g_dx; m.g_dy = in.g_dy;
    }
    m.bxmin = m.bymin = 1e300;
    m.bxmax = m.bymax = -1e300;
    for (int b = 0; b < in.nb; ++b) {
      m.bxmin = std::min(m.bxmin, in.xb[2 * b]); m.bxmax = std::max(m.bxmax, in.xb[2 * b]);
      m.bymin = std::min(m.bymin, in.xb[2 * b + 1]);
      m.bymax = std::max(m.bymax, in.xb[2 * b + 1]);
    }
    nids[nk + j] = next_id_++;
  }
  // offsets of the new layout
  long ntot = 0, nbtot = 0, ctot = 0, qtot = 0, gtot = 0, nblk = 0;
  for (int k = 0; k < ninst; ++k) {
    InstMeta &m = nm[k];
    m.p_off = (int)ntot; m.b_off = (int)nbtot; m.c_off = (int)ctot; m.q_off = (int)qtot;
    m.blk_off = (int)nblk; m.nblk = cdiv(m.n, kTPB);
    m.g_off = (int)gtot;
    ntot += m.n; nbtot += m.nb; ctot += m.c_cap; qtot += m.q_cap;
    nblk += m.nblk;
    if (m.wind_mode == 1) gtot += 2L * m.g_nx * m.g_ny;
  }
  if (ntot > INT32_MAX / kNbrMax || ctot > INT32_MAX - 2 || qtot > INT32_MAX) return -3;

  // new buffers
  DevState nd{};
  std::vector<void *> nlist;
  auto hip_alloc = [&nlist](void **p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess) {
      *p = nullptr;
      return 1;
    }
    nlist.push_back(*p);
    return 0;
  };
  int *nscan = nullptr;
  int rc = alloc_state(nd, hip_alloc, &nscan, ntot, nbtot, ctot, nblk, ninst, gtot) ? -6 : 0;
  if (rc == 0 && dalloc(hip_alloc, &nd.kz, ntot)) rc = -6;
  // the switches' arrays, when an instance uses them
  const bool x_sle = std::any_of(nm.begin(), nm.end(), [](const InstMeta &m) { return m.sle != 2; });
  const bool x_av = std::any_of(nm.begin(), nm.end(), [](const InstMeta &m) { return m.avis != 0; });
  const bool x_wang = std::any_of(nm.begin(), nm.end(), [](const InstMeta &m) { return m.bfa == 2; });
  const bool x_end = std::any_of(nm.begin(), nm.end(), [](const InstMeta &m) { return m.endp != 0; });
  // pair ranks of the viscosity's pi_ij: every pair of every instance has a
  // rank in the packed order; the viscosity needs the stored neighbour lists
  // (at most kNbrMax entries, each pair in two lists), so kNbrMax / 2 per
  // particle bounds the pairs
  const long pcap = x_av ? std::max<long>((long)(kNbrMax / 2) * ntot, kPairBlk) : 0;
  if (rc == 0 && x_sle) {
    if (dalloc(hip_alloc, &nd.drp, ntot) || dalloc(hip_alloc, &nd.masst[0], ntot) ||
        dalloc(hip_alloc, &nd.masst[1], ntot))
      rc = -6;
  }
  if (rc == 0 && x_end) {
    if (dalloc(hip_alloc, &nd.hnew, ntot) || dalloc(hip_alloc, &nd.cfl, ntot)) rc = -6;
  }
  if (rc == 0 && x_av) {
    if (dalloc(hip_alloc, &nd.lv, ntot) || dalloc(hip_alloc, &nd.pidx, (size_t)kNbrMax * ntot) ||
        dalloc(hip_alloc, &nd.piv, pcap) || dalloc(hip_alloc, &nd.pflg, pcap) ||
        dalloc(hip_alloc, &nd.pblast, (size_t)cdiv(pcap, kPairBlk)))
      rc = -6;
  }
  if (rc == 0 && x_wang) {
    if (dalloc(hip_alloc, &nd.bnx, nbtot) || dalloc(hip_alloc, &nd.bny, nbtot)) rc = -6;
  }
  nd.pcap = (int)pcap;
  if (rc == 0 && hipMalloc((void **)&nd.hq, std::max<long>(qtot, 1) * sizeof(double)) != hipSuccess)
    rc = -6;
  if (rc == 0 && dalloc(hip_alloc, &nd.hpair, qtot)) rc = -6;
  if (rc) {
    free_list(nlist);           // (scan_tmp is on the list)
    if (nd.hq) hipFree(nd.hq);
    return rc;
  }
  auto fail = [&](int code) {
    hipStreamSynchronize(st_);
    free_list(nlist);
    hipFree(nd.hq);
    return code;
  };
#define RCHK(x) \
  do {          \
    if ((x) != hipSuccess) return fail(-2); \
  } while (0)
  RCHK(hipMemsetAsync(nd.hq, 0, std::max<long>(qtot, 1) * sizeof(double), st_));
  RCHK(hipMemsetAsync(nd.count, 0, (ctot + 1) * sizeof(int), st_));
  RCHK(hipMemsetAsync(nd.fill, 0, (ctot + 1) * sizeof(int), st_));
  // no key yet: an added instance's first evaluation builds its lists
  RCHK(hipMemsetAsync(nd.cellid[0], 0xff, std::max<long>(ntot, 1) * sizeof(int), st_));
  // kernel_zero starts at zero, as the reference's module array does
  RCHK(hipMemsetAsync(nd.kz, 0, std::max<long>(ntot, 1) * sizeof(double), st_));
  // and so does Benz's drhodt (initialize.f90:44)
  if (nd.drp) RCHK(hipMemsetAsync(nd.drp, 0, std::max<long>(ntot, 1) * sizeof(double), st_));
  RCHK(hipMemsetAsync(nd.bflag, 0xff, (ctot + 1) * sizeof(int), st_));
  RCHK(hipMemsetAsync(nd.bcellc, 0xff, (ctot + 1) * sizeof(int), st_));
  // boundary lists are (re)built at the first evaluation of a layout
  RCHK(hipMemsetAsync(nd.bkey, 0xff, std::max<long>(nbtot, 1) * sizeof(int), st_));
  RCHK(hipMemsetAsync(nd.bcnt, 0, 9 * std::max<long>(nbtot, 1) * sizeof(int), st_));
  RCHK(hipMemsetAsync(nd.cellh, 0, (ctot + 1) * sizeof(unsigned long long), st_));
  RCHK(hipMemsetAsync(nd.osig, 0, std::max<long>(ntot, 1) * sizeof(int), st_));
  {
    // the first evaluation of a new layout builds the cell lists
    static const int one = 1;
    RCHK(hipMemcpyAsync(nd.rebuild, &one, sizeof(int), hipMemcpyHostToDevice, st_));
    RCHK(hipMemsetAsync(nd.bar, 0, kBarWords * sizeof(unsigned), st_));
  }

  // kept instances: device to device, from the set holding each one (its
  // cell lists and member counts stay valid: no rebuild unless a key changes)
  auto d2d = [&](void *dst, const void *src, size_t bytes) {
    return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st_) : hipSuccess;
  };
  std::vector<InstDyn> kdyn(nk);
  for (int j = 0; j < nk; ++j) {
    RCHK(hipMemcpyAsync(&kdyn[j], d_.dyn + keep[j], sizeof(InstDyn), hipMemcpyDeviceToHost, st_));
  }
  RCHK(hipStreamSynchronize(st_));
  for (int j = 0; j < nk; ++j) {
    const InstMeta &o = meta_[keep[j]];
    const InstMeta &m = nm[j];
    const size_t n8 = (size_t)m.n * 8, n4 = (size_t)m.n * 4;
    const int cs = kdyn[j].sel ^ kdyn[j].rb;
    kdyn[j].sel = 0;
    kdyn[j].rb = 0;
    kdyn[j].vbuilt = 0;     // candidate lists are built again
    kdyn[j].rchg = 1;       // and the pair-rank prefixes
    kdyn[j].vskin_next = 0.f;
    RCHK(d2d(nd.x[0] + m.p_off, d_.x[2 * cs] + o.p_off, n8));
    RCHK(d2d(nd.y[0] + m.p_off, d_.y[2 * cs] + o.p_off, n8));
    RCHK(d2d(nd.vx[0] + m.p_off, d_.vx[2 * cs] + o.p_off, n8));
    RCHK(d2d(nd.vy[0] + m.p_off, d_.vy[2 * cs] + o.p_off, n8));
    RCHK(d2d(nd.h[0] + m.p_off, d_.h[2 * cs] + o.p_off, n8));
    RCHK(d2d(nd.A[0] + m.p_off, d_.A[2 * cs] + o.p_off, n8));
    RCHK(d2d(nd.rho[0] + m.p_off, d_.rho[2 * cs] + o.p_off, n8));
    RCHK(d2d(nd.hsml[0] + m.p_off, d_.hsml[2 * cs] + o.p_off, n8));
    RCHK(d2d(nd.mass[0] + m.p_off, d_.mass[cs] + o.p_off, n8));
    RCHK(d2d(nd.kz + m.p_off, d_.kz + o.p_off, n8));
    RCHK(d2d(nd.orig[0] + m.p_off, d_.orig[cs] + o.p_off, n4));
    RCHK(d2d(nd.cellid[0] + m.p_off, d_.cellid[cs] + o.p_off, n4));
    RCHK(d2d(nd.count + m.c_off, d_.count + o.c_off, (size_t)m.c_cap * 4));
    RCHK(d2d(nd.xb + m.b_off, d_.xb + o.b_off, (size_t)m.nb * 8));
    RCHK(d2d(nd.yb + m.b_off, d_.yb + o.b_off, (size_t)m.nb * 8));
    RCHK(d2d(nd.massb + m.b_off, d_.massb + o.b_off, (size_t)m.nb * 8));
    RCHK(d2d(nd.hq + m.q_off, d_.hq + o.q_off, (size_t)m.q_cap * 8));
    if (m.sle == 1) RCHK(d2d(nd.drp + m.p_off, d_.drp + o.p_off, n8));
    if (m.bfa == 2) {
      RCHK(d2d(nd.bnx + m.b_off, d_.bnx + o.b_off, (size_t)m.nb * 8));
      RCHK(d2d(nd.bny + m.b_off, d_.bny + o.b_off, (size_t)m.nb * 8));
    }
    if (m.wind_mode == 1)
      RCHK(d2d(nd.grid + m.g_off, d_.grid + o.g_off, 16 * (size_t)m.g_nx * m.g_ny));
  }
  if (nk > 0)
    RCHK(hipMemcpyAsync(nd.dyn, kdyn.data(), sizeof(InstDyn) * nk, hipMemcpyHostToDevice, st_));
  if (na > 0 && carry_store_from >= 0 && carry_store_from < (int)meta_.size() && d_.hq) {
    const InstMeta &o = meta_[carry_store_from];
    RCHK(d2d(nd.hq + nm[nk].q_off, d_.hq + o.q_off, (size_t)std::min(o.q_cap, nm[nk].q_cap) * 8));
    // kernel_zero persists by particle index across calls as well, and so
    // does Benz's drhodt
    RCHK(d2d(nd.kz + nm[nk].p_off, d_.kz + o.p_off, (size_t)std::min(o.n, nm[nk].n) * 8));
    if (nd.drp && d_.drp)
      RCHK(d2d(nd.drp + nm[nk].p_off, d_.drp + o.p_off, (size_t)std::min(o.n, nm[nk].n) * 8));
  }

  // added instances: host to device (staged in SoA)
  std::vector<std::vector<double>> keepalive;
  std::vector<std::vector<int>> keepalive_i;
  auto h2d = [&](void *dst, const void *src, size_t bytes) {
    return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st_) : hipSuccess;
  };
  std::vector<InstDyn> ndyn(na);
  for (int j = 0; j < na; ++j) {
    const InstanceInput &in = add[j];
    const InstMeta &m = nm[nk + j];
    const int n = in.n;
    keepalive.emplace_back(12 * (size_t)n + 5 * (size_t)in.nb);
    std::vector<double> &st = keepalive.back();
    double *hx = st.data(), *hy = hx + n, *hvx = hy + n, *hvy = hvx + n;
    for (int i = 0; i < n; ++i) {
      hx[i] = in.x[2 * i]; hy[i] = in.x[2 * i + 1];
      hvx[i] = in.vx[2 * i]; hvy[i] = in.vx[2 * i + 1];
    }
    double *hxb = hvy + n, *hyb = hxb + in.nb, *hnx = hyb + in.nb, *hny = hnx + in.nb;
    for (int b = 0; b < in.nb; ++b) {
      hxb[b] = in.xb[2 * b];
      hyb[b] = in.xb[2 * b + 1];
      hnx[b] = in.bnormal ? in.bnormal[2 * b] : 0.0;
      hny[b] = in.bnormal ? in.bnormal[2 * b + 1] : 0.0;
    }
    keepalive_i.emplace_back(n);
    for (int i = 0; i < n; ++i) keepalive_i.back()[i] = i;
    const size_t n8 = (size_t)n * 8;
    RCHK(h2d(nd.x[0] + m.p_off, hx, n8));
    RCHK(h2d(nd.y[0] + m.p_off, hy, n8));
    RCHK(h2d(nd.vx[0] + m.p_off, hvx, n8));
    RCHK(h2d(nd.vy[0] + m.p_off, hvy, n8));
    RCHK(h2d(nd.h[0] + m.p_off, in.hth, n8));
    RCHK(h2d(nd.A[0] + m.p_off, in.A, n8));
    RCHK(h2d(nd.rho[0] + m.p_off, in.rho, n8));
    RCHK(h2d(nd.hsml[0] + m.p_off, in.hsml, n8));
    RCHK(h2d(nd.mass[0] + m.p_off, in.mass, n8));
    RCHK(h2d(nd.orig[0] + m.p_off, keepalive_i.back().data(), (size_t)n * 4));
    if (in.nb > 0) {
      RCHK(h2d(nd.xb + m.b_off, hxb, (size_t)in.nb * 8));
      RCHK(h2d(nd.yb + m.b_off, hyb, (size_t)in.nb * 8));
      RCHK(h2d(nd.massb + m.b_off, in.massb, (size_t)in.nb * 8));
      if (m.bfa == 2) {
        RCHK(h2d(nd.bnx + m.b_off, hnx, (size_t)in.nb * 8));
        RCHK(h2d(nd.bny + m.b_off, hny, (size_t)in.nb * 8));
      }
    }
    if (in.wind_mode == 1) {
      const size_t gg = (size_t)in.g_nx * in.g_ny;
      RCHK(h2d(nd.grid + m.g_off, in.g_u, gg * 8));
      RCHK(h2d(nd.grid + m.g_off + gg, in.g_v, gg * 8));
    }
    InstDyn &y = ndyn[j];
    std::memset(&y, 0, sizeof(y));
    y.dt = in.dt;
    y.time = in.time;
    y.itimestep = in.itimestep;
    y.active = 1;
    y.ipstep = -1;
    y.steps_left = in.nsteps > 0 ? in.nsteps : -1;
    y.rchg = 1;
  }
  if (na > 0) RCHK(h2d(nd.dyn + nk, ndyn.data(), sizeof(InstDyn) * na));
  RCHK(h2d(nd.meta, nm.data(), sizeof(InstMeta) * ninst));
  // Workgroups are dispatched to the 8 XCDs round-robin by block id; the
  // physical table gives each XCD a contiguous run of logical blocks.
  std::vector<BlockEntry> lg;
  for (int k = 0; k < ninst; ++k)
    for (int f = 0; f < nm[k].n; f += kTPB)
      lg.push_back(BlockEntry{k, f, (int)lg.size(), std::min(nm[k].n - f, kTPB),
                              nm[k].p_off + f, 0, nm[k].sle != 2 ? 1 : 0, 0});
  std::vector<BlockEntry> phys(lg.size());
  {
    const int nb_ = (int)lg.size(), q = nb_ / 8, r = nb_ % 8;
    for (int b = 0; b < nb_; ++b) {
      const int x = b % 8;
      phys[b] = lg[x * q + std::min(x, r) + b / 8];
    }
  }
  if (!phys.empty()) RCHK(h2d(nd.blk, phys.data(), sizeof(BlockEntry) * phys.size()));
  blk_host_ = phys;
  std::vector<int> cinst(ctot + 1, 0);
  for (int k = 0; k < ninst; ++k)
    std::fill(cinst.begin() + nm[k].c_off, cinst.begin() + nm[k].c_off + nm[k].c_cap, k);
  RCHK(h2d(nd.cinst, cinst.data(), sizeof(int) * cinst.size()));
  RCHK(hipStreamSynchronize(st_));
#undef RCHK

  // swap in the new layout
  graph_drop();
  free_list(allocs_);           // (scan_tmp_ is on the list)
  if (d_.hq && !arena_) hipFree(d_.hq);
  arena_ = false;
  arena_stage_ = nullptr;
  d_ = nd;
  d_.force_search = force_search_;
  d_.skin = tune_.skin;
  allocs_.swap(nlist);
  scan_tmp_ = nscan;
  d_.ntot = (int)ntot; d_.nbtot = (int)nbtot; d_.ctot = (int)ctot; d_.qtot = (int)qtot;
  d_.ninst = ninst;
  d_.nblk = (int)lg.size();
  meta_ = nm;
  ids_ = nids;
  any_boundary_ = nbtot > 0;
  // kernels with the switch values beyond the reference's defaults are
  // launched only when some instance uses one (each instance then follows
  // its own values)
  any_xsw_ = std::any_of(nm.begin(), nm.end(), [](const InstMeta &m) { return nondefault(m); });
  any_cden_ = std::any_of(nm.begin(), nm.end(), [](const InstMeta &m) { return m.den_alg == 2; });
  any_av_ = x_av;
  any_endp_ = x_end;
  if (new_ids) new_ids->assign(nids.begin() + nk, nids.end());
  if (ninst > 0) {
    // grid of the next evaluation of every instance
    launch_partial(d_, 0, st_);
    launch_finalize(d_, 0, st_);
    if (hipStreamSynchronize(st_) != hipSuccess) return -2;
  }
  return 0;
}

// Host loops over n items split over a few threads (the staging copies of
// the per-call drop-in are host-memory bound).
template <class F>
static void host_for(int n, F &&f) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int T = (int)std::min<unsigned>(std::min(hw, 8u), (unsigned)std::max(n / 16384, 1));
  if (T <= 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  const int per = (n + T - 1) / T;
  for (int t = 1; t < T; ++t)
    th.emplace_back([&, t] { f(std::min(n, t * per), std::min(n, (t + 1) * per)); });
  f(0, std::min(n, per));
  for (auto &x : th) x.join();
}

int Engine::stage_reserve(size_t bytes) {
  if (bytes <= stage_bytes_) return 0;
  if (stage_) hipHostFree(stage_);
  stage_ = nullptr;
  stage_bytes_ = 0;
  if (hipHostMalloc(&stage_, bytes, hipHostMallocDefault) != hipSuccess) {
    stage_ = nullptr;
    return -6;
  }
  stage_bytes_ = bytes;
  return 0;
}

int Engine::reload(const InstanceInput &in, bool grid_same) {
  if (meta_.size() != 1 || d_.ninst != 1 || blk_host_.empty()) return 1;
  const InstMeta &o = meta_[0];
  if (in.n != o.n || in.nb != o.nb || in.tse != o.tse || in.wind_mode != o.wind_mode ||
      in.den_alg != o.den_alg || in.pa_sph != o.pa_sph ||
      (in.pressure_force ? in.eos_alg : (in.viscosity ? 1 : 0)) != o.eos_alg ||
      in.rheology != o.rheology || in.skf != o.skf || in.sle != o.sle ||
      (in.xsph ? 1 : 0) != o.xsph || in.bfa != o.bfa ||
      ((in.xsph || in.art_visc) ? in.art_visc_alg : 0) != o.avis ||
      (in.art_visc || in.tse == 1 || in.boundary_friction ? 1 : 0) != o.needc ||
      (in.boundary_friction ? 1 : 0) != o.fric || (in.boundary_pressure ? 1 : 0) != o.bpres ||
      (in.wind_force ? 1 : 0) != o.wind_on || (in.water_force ? 1 : 0) != o.water_on ||
      (in.viscosity ? 0 : 1) != o.novisc)
    return 1;
  // (the boundary normals of the Wang force are laid out with the boundary)
  if (in.bfa == 2) return 1;
  if (in.wind_mode == 1 && (!grid_same || in.g_nx != o.g_nx || in.g_ny != o.g_ny)) return 1;
  if (!in.x || !in.vx || !in.mass || !in.rho || !in.hsml || !in.hth || !in.A) return -1;
  if (in.nb > 0 && (!in.xb || !in.massb)) return -1;
  const long cells = grid_cells(in.n, in.x, in.hsml, in.nb, in.xb, 1.0);
  if (cells < 0 || cells > (long)o.c_cap) return 1;
  HIPCHK(hipStreamSynchronize(st_));
  InstDyn y;
  HIPCHK(hipMemcpy(&y, d_.dyn, sizeof(y), hipMemcpyDeviceToHost));
  // the boundary-pair store must hold the new state's pairs: the last
  // evaluation's count with room to spare (a full upload counts them)
  if (in.nb > 0 && 3L * y.bnip / 2 + 512 > (long)o.q_cap) return 1;
  const int n = in.n, nb = in.nb;
  const size_t n8 = (size_t)n * 8, b8 = (size_t)nb * 8;
  int rc = stage_reserve(9 * n8 + 3 * b8 + (size_t)n * 4);
  if (rc) return rc;
  double *hx = (double *)stage_, *hy = hx + n, *hvx = hy + n, *hvy = hvx + n;
  double *hh = hvy + n, *hA = hh + n, *hr = hA + n, *hs = hr + n, *hm = hs + n;
  double *hxb = hm + n, *hyb = hxb + nb, *hmb = hyb + nb;
  int *horig = (int *)(hmb + nb);
  host_for(n, [&](int i0, int i1) {
    for (int i = i0; i < i1; ++i) {
      hx[i] = in.x[2 * i]; hy[i] = in.x[2 * i + 1];
      hvx[i] = in.vx[2 * i]; hvy[i] = in.vx[2 * i + 1];
      horig[i] = i;
    }
    const size_t l = (size_t)(i1 - i0) * 8;
    std::memcpy(hh + i0, in.hth + i0, l);
    std::memcpy(hA + i0, in.A + i0, l);
    std::memcpy(hr + i0, in.rho + i0, l);
    std::memcpy(hs + i0, in.hsml + i0, l);
    std::memcpy(hm + i0, in.mass + i0, l);
  });
  InstMeta m = o;
  m.wind_u = in.wind[0]; m.wind_v = in.wind[1];
  m.cur_u = in.current[0]; m.cur_v = in.current[1];
  m.max_hsml = in.max_hsml;
  m.bxmin = m.bymin = 1e300;
  m.bxmax = m.bymax = -1e300;
  for (int b = 0; b < nb; ++b) {
    hxb[b] = in.xb[2 * b]; hyb[b] = in.xb[2 * b + 1]; hmb[b] = in.massb[b];
    m.bxmin = std::min(m.bxmin, hxb[b]); m.bxmax = std::max(m.bxmax, hxb[b]);
    m.bymin = std::min(m.bymin, hyb[b]); m.bymax = std::max(m.bymax, hyb[b]);
  }
  // the instance record: a new state in set 0, lists to be built, the
  // evaluation stamps continuing (stamped cells of earlier evaluations stay
  // stale)
  y.dt = in.dt; y.time = in.time; y.itimestep = in.itimestep;
  y.err = 0; y.active = 1; y.steps_left = -1;
  y.rchg = 1; y.vbuilt = 0; y.vskin_next = 0.f; y.vmoved = 0.f; y.vprev = 0.f;
  y.sel = 0; y.rb = 0; y.ipstep = -1;
  auto h2d = [&](void *dst, const void *src, size_t bytes) {
    return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st_) : hipSuccess;
  };
  HIPCHK(h2d(d_.x[0], hx, n8)); HIPCHK(h2d(d_.y[0], hy, n8));
  HIPCHK(h2d(d_.vx[0], hvx, n8)); HIPCHK(h2d(d_.vy[0], hvy, n8));
  HIPCHK(h2d(d_.h[0], hh, n8)); HIPCHK(h2d(d_.A[0], hA, n8));
  HIPCHK(h2d(d_.rho[0], hr, n8)); HIPCHK(h2d(d_.hsml[0], hs, n8));
  HIPCHK(h2d(d_.mass[0], hm, n8));
  HIPCHK(h2d(d_.orig[0], horig, (size_t)n * 4));
  if (nb > 0) {
    HIPCHK(h2d(d_.xb, hxb, b8)); HIPCHK(h2d(d_.yb, hyb, b8)); HIPCHK(h2d(d_.massb, hmb, b8));
  }
  // no key yet: the first evaluation builds the lists
  HIPCHK(hipMemsetAsync(d_.cellid[0], 0xff, (size_t)n * 4, st_));
  HIPCHK(hipMemsetAsync(d_.count, 0, ((size_t)o.c_cap + 1) * 4, st_));
  HIPCHK(hipMemsetAsync(d_.fill, 0, ((size_t)o.c_cap + 1) * 4, st_));
  if (nb > 0) {
    HIPCHK(hipMemsetAsync(d_.bkey, 0xff, (size_t)nb * 4, st_));
    HIPCHK(hipMemsetAsync(d_.bcnt, 0, 9 * (size_t)nb * 4, st_));
  }
  static const int one = 1;
  HIPCHK(h2d(d_.rebuild, &one, sizeof(int)));
  HIPCHK(hipMemsetAsync(d_.bar, 0, kBarWords * sizeof(unsigned), st_));
  HIPCHK(h2d(d_.blk, blk_host_.data(), sizeof(BlockEntry) * blk_host_.size()));
  HIPCHK(h2d(d_.meta, &m, sizeof(m)));
  HIPCHK(h2d(d_.dyn, &y, sizeof(y)));
  meta_[0] = m;
  // grid of the first evaluation
  launch_partial(d_, 0, st_);
  launch_finalize(d_, 0, st_);
  HIPCHK(hipStreamSynchronize(st_));
  return 0;
}

int Engine::download_single(double *dxdt, double *dvxdt, double *dhdt, double *dAdt,
                            double *eps3, double *zeta, double *eta, double *p, double *sig3) {
  if (d_.ninst < 1) return -1;
  HIPCHK(hipStreamSynchronize(st_));
  const int n = meta_[0].n, N = d_.ntot, off = meta_[0].p_off;
  const int cs = cur_set(0);
  if (cs < 0) return -2;
  // 15 fields by slot and the slot order, in one batch
  const size_t n8 = (size_t)n * 8;
  int rc = stage_reserve(15 * n8 + (size_t)n * 4);
  if (rc) return rc;
  double *st = (double *)stage_;
  int *ord = (int *)(st + 15 * (size_t)n);
  auto d2h = [&](void *dst, const void *src, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st_);
  };
  HIPCHK(d2h(ord, d_.orig[cs] + off, (size_t)n * 4));
  const bool want_t = dxdt || dvxdt || dhdt || dAdt, want_d = eps3 || zeta || eta || p || sig3;
  for (int f = 0; f < 6 && want_t; ++f)
    HIPCHK(d2h(st + f * (size_t)n, d_.tend + (size_t)f * N + off, n8));
  for (int f = 0; f < 6 && want_d; ++f)
    HIPCHK(d2h(st + (6 + f) * (size_t)n, d_.diag + (size_t)f * N + off, n8));
  for (int f = 0; f < 3 && sig3; ++f)
    HIPCHK(d2h(st + (12 + f) * (size_t)n, d_.sig_out + (size_t)f * N + off, n8));
  HIPCHK(hipStreamSynchronize(st_));
  auto F = [&](int f, int i) { return st[f * (size_t)n + i]; };
  host_for(n, [&](int i0, int i1) {
    for (int i = i0; i < i1; ++i) {
      const int r = ord[i];
      if (dxdt) { dxdt[2 * r] = F(0, i); dxdt[2 * r + 1] = F(1, i); }
      if (dvxdt) { dvxdt[2 * r] = F(2, i); dvxdt[2 * r + 1] = F(3, i); }
      if (dhdt) dhdt[r] = F(4, i);
      if (dAdt) dAdt[r] = F(5, i);
      if (eps3) { eps3[3 * r] = F(6, i); eps3[3 * r + 1] = F(7, i); eps3[3 * r + 2] = F(8, i); }
      if (zeta) zeta[r] = F(9, i);
      if (eta) eta[r] = F(10, i);
      if (p) p[r] = F(11, i);
      if (sig3) { sig3[3 * r] = F(12, i); sig3[3 * r + 1] = F(13, i); sig3[3 * r + 2] = F(14, i); }
    }
  });
  return 0;
}

// ------------------------------------------- one instance in caller memory

namespace {
// Bump allocation from [base, base + cap) at 256-byte alignment; with a
// null base it only measures.
struct Carve {
  char *base;
  size_t cap, used = 0;
  int operator()(void **p, size_t bytes) {
    const size_t off = (used + 255) & ~(size_t)255;
    if (off > cap || bytes > cap - off) {
      *p = nullptr;
      return 1;
    }
    *p = base ? (void *)(base + off) : nullptr;
    used = off + bytes;
    return 0;
  }
};

// The part of the caller memory after the persistent arrays: the layout of
// one instance of `in` with c_cap cells, then the staging of the call's
// inputs and results (max(15 n, 4 n + 2 nb) doubles).
int carve_single(Carve &cv, DevState &nd, int **scan, double **stage, const InstanceInput &in,
                 long c_cap) {
  const size_t n = (size_t)in.n, nb = (size_t)in.nb;
  const size_t g = in.wind_mode == 1 ? 2 * (size_t)in.g_nx * in.g_ny : 0;
  const size_t nblk = (size_t)cdiv(in.n, kTPB);
  if (alloc_state(nd, cv, scan, n, nb, (size_t)c_cap, nblk, 1, g)) return 1;
  return dalloc(cv, stage, std::max(15 * n, 4 * n + 2 * nb));
}
}  // namespace

size_t Engine::single_bytes(const InstanceInput &in, long q_cap, long kz_cap) {
  const long cells = cell_capacity(in);
  if (cells < 0 || in.n < 1 || in.nb < 0) return 0;
  Carve cv{nullptr, ~(size_t)0};
  DevState nd{};
  int *scan = nullptr;
  double *stage = nullptr, *hq = nullptr, *kz = nullptr;
  int2 *hp = nullptr;
  dalloc(cv, &hq, (size_t)q_cap);
  dalloc(cv, &hp, (size_t)q_cap);
  dalloc(cv, &kz, (size_t)kz_cap);
  carve_single(cv, nd, &scan, &stage, in, cells);
  return cv.used;
}

int Engine::single_in_memory(char *base, size_t bytes, long q_cap, long kz_cap,
                             const InstanceInput &in, bool clear_persist, long kz_valid) {
  if (!meta_.empty() && !arena_) return -5;
  if (in.n < 1 || in.nb < 0 || in.nb > kMaxBoundaryPerInst) return -1;
  if (!in.x || !in.vx || !in.mass || !in.rho || !in.hsml || !in.hth || !in.A) return -1;
  if (in.nb > 0 && (!in.xb || !in.massb)) return -1;
  if (in.wind_mode == 1 && (in.g_nx < 2 || in.g_ny < 2 || !in.g_u || !in.g_v)) return -1;
  if (in.n > kz_cap || q_cap < 1 || q_cap > INT32_MAX) return -3;
  const long cells = cell_capacity(in);
  if (cells < 0) return -4;
  if (cells > (1L << 28) || (long)in.n > INT32_MAX / kNbrMax) return -3;
  // every pointer first: nothing is touched unless the whole layout fits
  Carve cv{base, bytes};
  DevState nd{};
  int *nscan = nullptr;
  double *stage = nullptr;
  if (dalloc(cv, &nd.hq, (size_t)q_cap) || dalloc(cv, &nd.hpair, (size_t)q_cap) ||
      dalloc(cv, &nd.kz, (size_t)kz_cap))
    return -3;
  if (carve_single(cv, nd, &nscan, &stage, in, cells)) return -3;

  InstMeta m;
  std::memset(&m, 0, sizeof(m));
  m.n = in.n;
  m.nb = in.nb;
  m.nblk = cdiv(in.n, kTPB);
  m.c_cap = (int)cells;
  m.q_cap = (int)q_cap;
  m.tse = in.tse;
  m.wind_mode = in.wind_mode;
  m.wind_u = in.wind[0]; m.wind_v = in.wind[1];
  m.cur_u = in.current[0]; m.cur_v = in.current[1];
  m.max_hsml = in.max_hsml;
  meta_switches(m, in);
  if (in.wind_mode == 1) {
    m.g_nx = in.g_nx; m.g_ny = in.g_ny;
    m.g_x0 = in.g_x0; m.g_y0 = in.g_y0; m.g_dx = in.g_dx; m.g_dy = in.g_dy;
  }
  m.bxmin = m.bymin = 1e300;
  m.bxmax = m.bymax = -1e300;
  for (int b = 0; b < in.nb; ++b) {
    m.bxmin = std::min(m.bxmin, in.xb[2 * b]); m.bxmax = std::max(m.bxmax, in.xb[2 * b]);
    m.bymin = std::min(m.bymin, in.xb[2 * b + 1]);
    m.bymax = std::max(m.bymax, in.xb[2 * b + 1]);
  }
  InstDyn y;
  std::memset(&y, 0, sizeof(y));
  y.dt = in.dt;
  y.time = in.time;
  y.itimestep = in.itimestep;
  y.active = 1;
  y.ipstep = -1;
  y.steps_left = -1;
  y.rchg = 1;
  nd.ntot = in.n; nd.nbtot = in.nb; nd.ctot = (int)cells; nd.qtot = (int)q_cap;
  nd.ninst = 1;
  nd.nblk = m.nblk;

  const size_t n = (size_t)in.n, nb = (size_t)in.nb, C1 = (size_t)cells + 1;
  auto h2d = [&](void *dst, const void *src, size_t b) {
    return b ? hipMemcpyAsync(dst, src, b, hipMemcpyHostToDevice, st_) : hipSuccess;
  };
  auto zero = [&](void *dst, int v, size_t b) {
    return b ? hipMemsetAsync(dst, v, b, st_) : hipSuccess;
  };
  // the persistent arrays: cleared when the caller starts them afresh,
  // otherwise kernel_zero of particles beyond the last call's count starts
  // at zero (the hsml_ij store keeps every entry)
  if (clear_persist) {
    HIPCHK(zero(nd.hq, 0, (size_t)q_cap * 8));
    HIPCHK(zero(nd.kz, 0, (size_t)kz_cap * 8));
  } else if ((long)n > kz_valid) {
    const long from = std::max(kz_valid, 0L);
    HIPCHK(zero(nd.kz + from, 0, (size_t)(n - from) * 8));
  }
  // what a new layout starts from (as a packed layout does)
  HIPCHK(zero(nd.count, 0, C1 * 4));
  HIPCHK(zero(nd.fill, 0, C1 * 4));
  HIPCHK(zero(nd.cinst, 0, C1 * 4));
  HIPCHK(zero(nd.cellid[0], 0xff, n * 4));
  HIPCHK(zero(nd.bflag, 0xff, C1 * 4));
  HIPCHK(zero(nd.bcellc, 0xff, C1 * 4));
  HIPCHK(zero(nd.bkey, 0xff, std::max<size_t>(nb, 1) * 4));
  HIPCHK(zero(nd.bcnt, 0, 9 * std::max<size_t>(nb, 1) * 4));
  HIPCHK(zero(nd.cellh, 0, C1 * 8));
  HIPCHK(zero(nd.osig, 0, n * 4));
  static const int one = 1;
  HIPCHK(h2d(nd.rebuild, &one, sizeof(int)));
  HIPCHK(zero(nd.bar, 0, kBarWords * sizeof(unsigned)));
  // the state: (2, n) arrays through the staging area, the others directly
  double *sx = stage, *sv = stage + 2 * n, *sb = stage + 4 * n;
  HIPCHK(h2d(sx, in.x, 2 * n * 8));
  HIPCHK(h2d(sv, in.vx, 2 * n * 8));
  HIPCHK(h2d(sb, in.xb, 2 * nb * 8));
  HIPCHK(h2d(nd.h[0], in.hth, n * 8));
  HIPCHK(h2d(nd.A[0], in.A, n * 8));
  HIPCHK(h2d(nd.rho[0], in.rho, n * 8));
  HIPCHK(h2d(nd.hsml[0], in.hsml, n * 8));
  HIPCHK(h2d(nd.mass[0], in.mass, n * 8));
  HIPCHK(h2d(nd.massb, in.massb, nb * 8));
  if (in.wind_mode == 1) {
    const size_t gg = (size_t)in.g_nx * in.g_ny;
    HIPCHK(h2d(nd.grid, in.g_u, gg * 8));
    HIPCHK(h2d(nd.grid + gg, in.g_v, gg * 8));
  }
  HIPCHK(h2d(nd.meta, &m, sizeof(m)));
  HIPCHK(h2d(nd.dyn, &y, sizeof(y)));
  launch_single_unpack(nd, sx, sv, sb, in.n, in.nb, st_);
  launch_single_blocks(nd, in.n, st_);

  d_ = nd;
  d_.force_search = force_search_;
  d_.skin = tune_.skin;
  scan_tmp_ = nscan;
  arena_ = true;
  arena_stage_ = stage;
  blk_host_.clear();
  meta_.assign(1, m);
  ids_.assign(1, next_id_++);
  any_boundary_ = in.nb > 0;
  any_xsw_ = nondefault(m);
  any_cden_ = m.den_alg == 2;
  any_av_ = false;      // (the caller rejects the switches that need arrays
  any_endp_ = false;    //  this layout does not carve)
  // grid of the first evaluation
  launch_partial(d_, 0, st_);
  launch_finalize(d_, 0, st_);
  HIPCHK(hipStreamSynchronize(st_));
  return 0;
}

int Engine::single_in_memory_out(double *dxdt, double *dvxdt, double *dhdt, double *dAdt,
                                 double *eps3, double *zeta, double *eta, double *p,
                                 double *sig3) {
  if (!arena_ || d_.ninst != 1) return -5;
  const size_t n = (size_t)meta_[0].n;
  const bool want_t = dxdt || dvxdt || dhdt || dAdt, want_d = eps3 || zeta || eta || p || sig3;
  launch_single_pack(d_, arena_stage_, (int)n, want_t, want_d, st_);
  auto d2h = [&](double *dst, size_t off, size_t cnt) {
    return dst ? hipMemcpyAsync(dst, arena_stage_ + off * n, cnt * n * 8, hipMemcpyDeviceToHost,
                                st_)
               : hipSuccess;
  };
  HIPCHK(d2h(dxdt, 0, 2));
  HIPCHK(d2h(dvxdt, 2, 2));
  HIPCHK(d2h(dhdt, 4, 1));
  HIPCHK(d2h(dAdt, 5, 1));
  HIPCHK(d2h(eps3, 6, 3));
  HIPCHK(d2h(zeta, 9, 1));
  HIPCHK(d2h(eta, 10, 1));
  HIPCHK(d2h(p, 11, 1));
  HIPCHK(d2h(sig3, 12, 3));
  HIPCHK(hipStreamSynchronize(st_));
  return 0;
}

template <class F>
void Engine::L(const char *name, F &&f) {
  if (!prof) {
    f();
    return;
  }
  KernelEvents ke;
  g_kernel_events = &ke;
  f();
  g_kernel_events = nullptr;
  Profiler::Rec r;
  r.name = name;
  r.ev = std::move(ke.ev);
  prof->recs.push_back(std::move(r));
}

void Engine::enqueue_prefix(int buf) {
  const DevState &d = d_;
  L("count", [&] { launch_count(d, buf, st_); });
  L("rebuild", [&] { launch_rebuild(d, buf, scan_tmp_, st_); });
}

void Engine::enqueue_eval(int buf, int phase, bool want_diag) {
  const DevState &d = d_;
  enqueue_prefix(buf);
  L("passA", [&] {
    launch_passA(d, buf, precision_, any_xsw_, want_diag ? 1 : 0, tune_.passA_batch, st_);
  });
  if (phase == 2 || any_av_) {
    // all pair ranks: nip is reported, and the artificial viscosity's pi_ij
    // runs over the pairs in their order
    L("pairtot", [&] { launch_pairtot(d, st_); });
    L("scan", [&] { launch_scan(d.pair_count, d.ctot, scan_tmp_, d.pair_start, 0, nullptr, st_); });
  }
  if (any_av_) L("avpairs", [&] { launch_av(d, buf, st_); });
  // boundary-pair ranks and the head of the pair ranks the boundary kernel reads
  if (any_boundary_) L("ranks", [&] { launch_ranks(d, buf, st_); });
  L("passB", [&] { launch_passB(d, precision_, any_xsw_, phase, tune_.passB_batch, st_); });
}

void Engine::enqueue_step() {
  if (scheme_ == 1) {
    // euler_scheme: one evaluation of the step's state, then the update
    enqueue_eval(0, 3, diag_);
    if (any_cden_) L("condens", [&] { launch_condens(d_, 1, st_); });
    if (any_endp_) L("endpass", [&] { launch_endpass(d_, 1, st_); });
    L("euler", [&] { launch_euler(d_, st_); });
    L("finalize", [&] { launch_finalize(d_, 1, st_); });
    return;
  }
  // predictor_corrector and leapfrog_scheme share the half step
  enqueue_eval(0, 0, false);
  if (any_cden_) L("condens", [&] { launch_condens(d_, 0, st_); });
  L("finalize", [&] { launch_finalize(d_, 0, st_); });
  enqueue_eval(1, scheme_ == 3 ? 4 : 1, diag_);
  if (any_cden_ && scheme_ == 3) L("condens", [&] { launch_condens(d_, 2, st_); });
  if (any_endp_) L("endpass", [&] { launch_endpass(d_, 0, st_); });
  L("finalize", [&] { launch_finalize(d_, 1, st_); });
}

void Engine::set_tuning(const Tuning &t) {
  tune_ = t;
  d_.skin = t.skin;
  graph_drop();           // the captured launches hold the old variants and values
}

int Engine::autotune(int nsteps, double *us_start, double *us_best) {
  if (nsteps < 1) return -1;
  if (d_.ninst == 0) return -5;
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  auto trial = [&](const Tuning &t, double *us) {
    set_tuning(t);
    int rc = this->steps(1, true);        // builds the step graph
    if (rc) return rc;
    HIPCHK(hipEventRecord(a, st_));
    rc = this->steps(nsteps, true);
    if (rc) return rc;
    HIPCHK(hipEventRecord(b, st_));
    HIPCHK(hipEventSynchronize(b));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, a, b));
    *us = 1e3 * (double)ms / nsteps;
    return 0;
  };
  Tuning best = tune_;
  double tb = 0.;
  int rc = trial(best, &tb);
  if (rc == 0 && us_start) *us_start = tb;
  auto consider = [&](Tuning t) {
    double us = 0.;
    if (rc == 0) rc = trial(t, &us);
    if (rc == 0 && us < tb) {
      tb = us;
      best = t;
    }
  };
  for (int ba : {1, 2, 3, 4}) {
    if (ba == best.passA_batch) continue;
    Tuning t = best;
    t.passA_batch = ba;
    consider(t);
  }
  for (int bb : {2, 3, 4}) {
    if (bb == best.passB_batch) continue;
    Tuning t = best;
    t.passB_batch = bb;
    consider(t);
  }
  for (float sk : {0.005f, 0.01f, 0.02f, 0.04f}) {
    if (sk == best.skin) continue;
    Tuning t = best;
    t.skin = sk;
    consider(t);
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  set_tuning(best);
  if (us_best) *us_best = tb;
  return rc;
}

void Engine::set_scheme(int int_algorithm) {
  if (scheme_ == int_algorithm) return;
  scheme_ = int_algorithm;
  graph_drop();
}

int Engine::eval_single(bool want_diag) {
  launch_stats_reset(d_, st_);
  enqueue_eval(0, 2, want_diag);
  launch_fold(d_, st_);
  return sync_and_check();
}

int Engine::graph_build() {
  if (gexec_) return 0;
  if (d_.ninst == 0) return 0;
  HIPCHK(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
  enqueue_step();
  HIPCHK(hipStreamEndCapture(st_, &graph_));
  HIPCHK(hipGraphInstantiate(&gexec_, graph_, nullptr, nullptr, 0));
  return 0;
}

int Engine::steps(int nsteps, bool use_graph) {
  if (nsteps < 0) return -1;
  if (d_.ninst == 0) return 0;
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
  if (d_.ninst == 1) {
    InstDyn y;
    HIPCHK(hipMemcpy(&y, d_.dyn, sizeof(y), hipMemcpyDeviceToHost));
    if (y.err & kErrSync) return -2;
    return (y.err & kErrNan) ? -4 : (y.err ? -3 : 0);
  }
  std::vector<InstDyn> hd(d_.ninst);
  HIPCHK(hipMemcpy(hd.data(), d_.dyn, sizeof(InstDyn) * d_.ninst, hipMemcpyDeviceToHost));
  int worst = 0;
  for (auto &y : hd) {
    if (y.err & kErrSync) return -2;
    if (y.err & kErrNan) return -4;
    if (y.err) worst = -3;
  }
  return worst;
}

void Engine::set_force_search(bool on) {
  if (force_search_ == on) return;
  force_search_ = on;
  d_.force_search = on;
  graph_drop();           // the captured launches hold the old value
}

int Engine::search_stats(int k, long long *evals, long long *cell, long long *cand) {
  if (k < 0 || k >= d_.ninst) return -1;
  HIPCHK(hipStreamSynchronize(st_));
  InstDyn y;
  HIPCHK(hipMemcpy(&y, d_.dyn + k, sizeof(y), hipMemcpyDeviceToHost));
  if (evals) *evals = (long long)y.n_eval;
  if (cell) *cell = (long long)y.n_cell;
  if (cand) *cand = (long long)y.n_cand;
  return 0;
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

int Engine::download_status(int k, int *itimestep, int *steps_left, int *active, int *err) {
  if (k < 0 || k >= d_.ninst) return -1;
  InstDyn y;
  HIPCHK(hipMemcpy(&y, d_.dyn + k, sizeof(y), hipMemcpyDeviceToHost));
  if (itimestep) *itimestep = y.itimestep;
  if (steps_left) *steps_left = y.steps_left;
  if (active) *active = y.active;
  if (err) *err = y.err;
  return 0;
}

int Engine::interaction_stats(int k, int *maxip, int *minip, long long *sumip, int *noip,
                              long long *nip, int *step) {
  if (k < 0 || k >= d_.ninst) return -1;
  HIPCHK(hipStreamSynchronize(st_));
  InstDyn y;
  HIPCHK(hipMemcpy(&y, d_.dyn + k, sizeof(y), hipMemcpyDeviceToHost));
  if (y.ipstep < 0 || y.ipmin == 0x7fffffff) return -5;
  if (maxip) *maxip = y.ipmax;
  if (minip) *minip = y.ipmin;
  if (sumip) *sumip = (long long)y.ipsum;
  if (noip) *noip = y.ipzero;
  if (nip) *nip = (long long)y.ipown;
  if (step) *step = y.ipstep;
  return 0;
}

int Engine::download_all_status(std::vector<InstDyn> &out) {
  out.resize(d_.ninst);
  if (d_.ninst == 0) return 0;
  HIPCHK(hipStreamSynchronize(st_));
  HIPCHK(hipMemcpy(out.data(), d_.dyn, sizeof(InstDyn) * d_.ninst, hipMemcpyDeviceToHost));
  return 0;
}

// Per-slot arrays come back in reference order through the reference index
// of each slot (orig).
struct SlotOrder {
  std::vector<int> ref;
  int fetch(const int *orig, int off, int n) {
    ref.resize(n);
    return hipMemcpy(ref.data(), orig + off, (size_t)n * 4, hipMemcpyDeviceToHost) == hipSuccess
               ? 0 : -2;
  }
  void put(const std::vector<double> &v, double *out) const {
    for (size_t i = 0; i < ref.size(); ++i) out[ref[i]] = v[i];
  }
  void put_pairs(const std::vector<double> &a, const std::vector<double> &b, double *out) const {
    for (size_t i = 0; i < ref.size(); ++i) {
      out[2 * ref[i]] = a[i];
      out[2 * ref[i] + 1] = b[i];
    }
  }
  void put_triples(const std::vector<double> &a, const std::vector<double> &b,
                   const std::vector<double> &c, double *out) const {
    for (size_t i = 0; i < ref.size(); ++i) {
      out[3 * ref[i]] = a[i];
      out[3 * ref[i] + 1] = b[i];
      out[3 * ref[i] + 2] = c[i];
    }
  }
};

// The set holding instance k's state (-1: the device could not be read).
int Engine::cur_set(int k) {
  InstDyn y;
  if (hipMemcpy(&y, d_.dyn + k, sizeof(y), hipMemcpyDeviceToHost) != hipSuccess) return -1;
  return y.sel ^ y.rb;
}

int Engine::download_state(int k, double *x, double *vx, double *rho, double *hsml, double *hth,
                           double *A) {
  if (k < 0 || k >= d_.ninst) return -1;
  HIPCHK(hipStreamSynchronize(st_));
  const InstMeta &m = meta_[k];
  const int n = m.n;
  const int cs = cur_set(k);
  if (cs < 0) return -2;
  SlotOrder so;
  if (so.fetch(d_.orig[cs], m.p_off, n)) return -2;
  std::vector<double> a(n), b(n);
  auto get = [&](std::vector<double> &dst, const double *src) {
    return hipMemcpy(dst.data(), src + m.p_off, (size_t)n * 8, hipMemcpyDeviceToHost);
  };
  if (x) {
    HIPCHK(get(a, d_.x[2 * cs]));
    HIPCHK(get(b, d_.y[2 * cs]));
    so.put_pairs(a, b, x);
  }
  if (vx) {
    HIPCHK(get(a, d_.vx[2 * cs]));
    HIPCHK(get(b, d_.vy[2 * cs]));
    so.put_pairs(a, b, vx);
  }
  const std::pair<double *, const double *> one[] = {{rho, d_.rho[2 * cs]},
                                                     {hsml, d_.hsml[2 * cs]},
                                                     {hth, d_.h[2 * cs]},
                                                     {A, d_.A[2 * cs]}};
  for (const auto &f : one) {
    if (!f.first) continue;
    HIPCHK(get(a, f.second));
    so.put(a, f.first);
  }
  return 0;
}

int Engine::download_tend(int k, double *dxdt, double *dvxdt, double *dhdt, double *dAdt) {
  if (k < 0 || k >= d_.ninst) return -1;
  HIPCHK(hipStreamSynchronize(st_));
  const InstMeta &m = meta_[k];
  const int n = m.n, N = d_.ntot;
  const int cs = cur_set(k);
  if (cs < 0) return -2;
  SlotOrder so;
  if (so.fetch(d_.orig[cs], m.p_off, n)) return -2;
  std::vector<double> a(n), b(n);
  auto get = [&](std::vector<double> &dst, int comp) {
    return hipMemcpy(dst.data(), d_.tend + (size_t)comp * N + m.p_off, (size_t)n * 8,
                     hipMemcpyDeviceToHost);
  };
  if (dxdt) {
    HIPCHK(get(a, 0));
    HIPCHK(get(b, 1));
    so.put_pairs(a, b, dxdt);
  }
  if (dvxdt) {
    HIPCHK(get(a, 2));
    HIPCHK(get(b, 3));
    so.put_pairs(a, b, dvxdt);
  }
  if (dhdt) {
    HIPCHK(get(a, 4));
    so.put(a, dhdt);
  }
  if (dAdt) {
    HIPCHK(get(a, 5));
    so.put(a, dAdt);
  }
  return 0;
}

int Engine::download_diag(int k, double *eps3, double *zeta, double *eta, double *p,
                          double *sig3) {
  if (k < 0 || k >= d_.ninst) return -1;
  HIPCHK(hipStreamSynchronize(st_));
  const InstMeta &m = meta_[k];
  const int n = m.n, N = d_.ntot;
  const int cs = cur_set(k);
  if (cs < 0) return -2;
  SlotOrder so;
  if (so.fetch(d_.orig[cs], m.p_off, n)) return -2;
  std::vector<double> c0(n), c1(n), c2(n);
  auto get = [&](std::vector<double> &dst, const double *src, int comp) {
    return hipMemcpy(dst.data(), src + (size_t)comp * N + m.p_off, (size_t)n * 8,
                     hipMemcpyDeviceToHost);
  };
  if (eps3) {
    HIPCHK(get(c0, d_.diag, 0));
    HIPCHK(get(c1, d_.diag, 1));
    HIPCHK(get(c2, d_.diag, 2));
    so.put_triples(c0, c1, c2, eps3);
  }
  const std::pair<double *, int> one[] = {{zeta, 3}, {eta, 4}, {p, 5}};
  for (const auto &f : one) {
    if (!f.first) continue;
    HIPCHK(get(c0, d_.diag, f.second));
    so.put(c0, f.first);
  }
  if (sig3) {
    HIPCHK(get(c0, d_.sig_out, 0));
    HIPCHK(get(c1, d_.sig_out, 1));
    HIPCHK(get(c2, d_.sig_out, 2));
    so.put_triples(c0, c1, c2, sig3);
  }
  return 0;
}

long Engine::store_cap(int k) const {
  return (k < 0 || k >= (int)meta_.size()) ? -1 : (long)meta_[k].q_cap;
}

int Engine::download_persist(int k, double *mass, double *xb, double *massb, double *kz,
                             double *hq, double *drp, double *bnormal) {
  if (k < 0 || k >= d_.ninst) return -1;
  HIPCHK(hipStreamSynchronize(st_));
  const InstMeta &m = meta_[k];
  const int cs = cur_set(k);
  if (cs < 0) return -2;
  if (mass) {
    SlotOrder so;
    if (so.fetch(d_.orig[cs], m.p_off, m.n)) return -2;
    std::vector<double> a(m.n);
    HIPCHK(hipMemcpy(a.data(), d_.mass[cs] + m.p_off, (size_t)m.n * 8, hipMemcpyDeviceToHost));
    so.put(a, mass);
  }
  if (m.nb > 0 && (xb || massb)) {
    std::vector<double> bx(m.nb), by(m.nb);
    HIPCHK(hipMemcpy(bx.data(), d_.xb + m.b_off, (size_t)m.nb * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(by.data(), d_.yb + m.b_off, (size_t)m.nb * 8, hipMemcpyDeviceToHost));
    if (xb)
      for (int b = 0; b < m.nb; ++b) {
        xb[2 * b] = bx[b];
        xb[2 * b + 1] = by[b];
      }
    if (massb)
      HIPCHK(hipMemcpy(massb, d_.massb + m.b_off, (size_t)m.nb * 8, hipMemcpyDeviceToHost));
  }
  if (kz) HIPCHK(hipMemcpy(kz, d_.kz + m.p_off, (size_t)m.n * 8, hipMemcpyDeviceToHost));
  if (hq && m.q_cap > 0)
    HIPCHK(hipMemcpy(hq, d_.hq + m.q_off, (size_t)m.q_cap * 8, hipMemcpyDeviceToHost));
  if (drp) {
    if (!d_.drp) return -5;
    HIPCHK(hipMemcpy(drp, d_.drp + m.p_off, (size_t)m.n * 8, hipMemcpyDeviceToHost));
  }
  if (bnormal && m.nb > 0) {
    if (!d_.bnx) return -5;
    std::vector<double> nx(m.nb), ny(m.nb);
    HIPCHK(hipMemcpy(nx.data(), d_.bnx + m.b_off, (size_t)m.nb * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ny.data(), d_.bny + m.b_off, (size_t)m.nb * 8, hipMemcpyDeviceToHost));
    for (int b = 0; b < m.nb; ++b) {
      bnormal[2 * b] = nx[b];
      bnormal[2 * b + 1] = ny[b];
    }
  }
  return 0;
}

int Engine::upload_persist(int k, const double *kz, const double *hq, long nq,
                           const double *drp) {
  if (k < 0 || k >= d_.ninst) return -1;
  const InstMeta &m = meta_[k];
  if (kz) HIPCHK(hipMemcpyAsync(d_.kz + m.p_off, kz, (size_t)m.n * 8, hipMemcpyHostToDevice, st_));
  if (drp) {
    if (!d_.drp) return -5;
    HIPCHK(hipMemcpyAsync(d_.drp + m.p_off, drp, (size_t)m.n * 8, hipMemcpyHostToDevice, st_));
  }
  const long q = std::min<long>(nq, m.q_cap);
  if (hq && q > 0)
    HIPCHK(hipMemcpyAsync(d_.hq + m.q_off, hq, (size_t)q * 8, hipMemcpyHostToDevice, st_));
  HIPCHK(hipStreamSynchronize(st_));
  return 0;
}

int Engine::download_kernel_zero(int k, double *kz) {
  if (k < 0 || k >= d_.ninst) return -1;
  HIPCHK(hipStreamSynchronize(st_));
  const InstMeta &m = meta_[k];
  HIPCHK(hipMemcpy(kz, d_.kz + m.p_off, (size_t)m.n * 8, hipMemcpyDeviceToHost));
  return 0;
}

int Engine::export_pairs(int cap, int *nip, int *pi, int *pj, double *rij, double *hij,
                         int bcap, int *bnip, int *bpi, int *bpj, double *brij, double *w,
                         double *dwdx) {
  // requires a prefix + passA (own counts) + pair ranks for state S
  const DevState &d = d_;
  enqueue_prefix(0);
  launch_passA(d, 0, precision_, any_xsw_, 0, tune_.passA_batch, st_);
  if (any_boundary_) launch_ranks(d, 0, st_);
  launch_pairtot(d, st_);
  launch_scan(d.pair_count, d.ctot, scan_tmp_, d.pair_start, 0, nullptr, st_);
  launch_fold(d, st_);
  int rc = sync_and_check();
  if (rc) return rc;
  int n_ip = 0, n_bp = 0, err = 0;
  rc = download_scalars(0, nullptr, nullptr, nullptr, &n_ip, &n_bp, &err);
  if (rc) return rc;
  if (!any_boundary_) n_bp = 0;
  *nip = n_ip;
  *bnip = n_bp;
  int *dpi = nullptr, *dpj = nullptr, *dbi = nullptr, *dbj = nullptr;
  double *dr = nullptr, *dh = nullptr, *dbr = nullptr, *dw = nullptr, *dg = nullptr;
  const int c1 = std::max(std::min(cap, n_ip), 1), c2 = std::max(std::min(bcap, n_bp), 1);
  bool ok = hipMalloc((void **)&dpi, c1 * 4) == hipSuccess &&
            hipMalloc((void **)&dpj, c1 * 4) == hipSuccess &&
            hipMalloc((void **)&dr, c1 * 8) == hipSuccess &&
            hipMalloc((void **)&dh, c1 * 8) == hipSuccess &&
            (!w || hipMalloc((void **)&dw, c1 * 8) == hipSuccess) &&
            (!dwdx || hipMalloc((void **)&dg, c1 * 16) == hipSuccess) &&
            hipMalloc((void **)&dbi, c2 * 4) == hipSuccess &&
            hipMalloc((void **)&dbj, c2 * 4) == hipSuccess &&
            hipMalloc((void **)&dbr, c2 * 8) == hipSuccess;
  if (ok) {
    launch_export_pairs(d, 0, std::min(cap, n_ip), dpi, dpj, dr, dh, dw, dg, st_);
    launch_export_bpairs(d, 0, any_boundary_ ? meta_[0].nb : 0, std::min(bcap, n_bp), dbi, dbj,
                         dbr, st_);
    ok = hipStreamSynchronize(st_) == hipSuccess;
    int k1 = std::min(cap, n_ip), k2 = std::min(bcap, n_bp);
    if (ok && k1 > 0) {
      if (pi) ok &= hipMemcpy(pi, dpi, k1 * 4, hipMemcpyDeviceToHost) == hipSuccess;
      if (pj) ok &= hipMemcpy(pj, dpj, k1 * 4, hipMemcpyDeviceToHost) == hipSuccess;
      if (rij) ok &= hipMemcpy(rij, dr, k1 * 8, hipMemcpyDeviceToHost) == hipSuccess;
      if (hij) ok &= hipMemcpy(hij, dh, k1 * 8, hipMemcpyDeviceToHost) == hipSuccess;
      if (w) ok &= hipMemcpy(w, dw, k1 * 8, hipMemcpyDeviceToHost) == hipSuccess;
      if (dwdx) ok &= hipMemcpy(dwdx, dg, k1 * 16, hipMemcpyDeviceToHost) == hipSuccess;
    }
    if (ok && k2 > 0) {
      if (bpi) ok &= hipMemcpy(bpi, dbi, k2 * 4, hipMemcpyDeviceToHost) == hipSuccess;
      if (bpj) ok &= hipMemcpy(bpj, dbj, k2 * 4, hipMemcpyDeviceToHost) == hipSuccess;
      if (brij) ok &= hipMemcpy(brij, dbr, k2 * 8, hipMemcpyDeviceToHost) == hipSuccess;
    }
  }
  hipFree(dpi); hipFree(dpj); hipFree(dr); hipFree(dh); hipFree(dw); hipFree(dg);
  hipFree(dbi); hipFree(dbj); hipFree(dbr);
  if (!ok) return -2;
  if (n_ip > cap || n_bp > bcap) return -3;
  return 0;
}

}  // namespace simp
