// simp_kernels.hip -- gfx950 kernels of the SPH viscous-plastic step.
//
// One right-hand-side evaluation (the work of SIMP's single_step,
// integration/single_step.f90:5-108) is a short pipeline of kernels over the
// packed particles of every instance:
//
//   finalize     grid of the evaluation from per-block partial reductions of
//                the bounding box and MINVAL(hsml) (nnps.f90:211-224); at
//                the end of a step also the VP dt, time and step counter
//                (time_step.f90:155, time_integration.f90:68-71); and the
//                per-cell lists of the boundary particles whose 3x3 block
//                covers the cell (nnps.f90:451-493), rebuilt only when a
//                boundary particle's cell changes
//   count        cell key NINT(x/bin) (nnps.f90:246-254); flags a rebuild
//                of the cell lists when any key changed
//   rebuild      the cell lists, members in ascending particle id (the
//                reference's bin order), and the state moved to the new
//                list order, in one persistent launch with grid barriers;
//                it returns at once when no key changed (the corrector
//                evaluation normally reuses the predictor's lists)
//   passA        neighbour gather: Wendland C6 gradient, velocity divergence
//                and gradient, strain rate and the VP rheology epilogue
//                (kernel.f90:471-520, vector_divergence.f90:27-28,
//                vector_gradient.f90:27-40, single_step.f90:80-84,
//                internal_stress.f90:21-53 and its callees); also the
//                boundary-pair counts per (boundary particle, cell)
//   ranks        boundary-pair ranks and bnip; hsml_ij of the particle
//                pairs ranked below bnip (read by b_Wendland_C6,
//                boundary_kernel.f90:491)
//   (pairtot, scan: all pair ranks, for the reported nip of single_step)
//   passB        neighbour gather: stress divergence (internal_stress.f90:
//                103-136), boundary force (boundary_kernel.f90:477-502,
//                boundary_force.f90:78-82), wind and water drag, tendencies
//                (single_step.f90:104-106), and the predictor or corrector
//                update (time_integration.f90:207-280); its blocks leave
//                partial min/max of the new positions and smoothing lengths
//                for the next finalize
//
// Every particle gathers its neighbours itself (no atomics on sums), visiting
// the 3x3 cell block in the order in which the reference's single-thread
// pair list reaches that particle: cells (-1,-1) (0,-1) (+1,-1) (-1,0), its
// own cell, (+1,0) (0,+1) (+1,+1) (-1,+1), members in ascending id.  With the
// per-pair arithmetic written in the reference's evaluation order and FP
// contraction off, each FP64 sum is formed in the same order as the
// reference's, so the FP64 path reproduces the reference's results.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <type_traits>

#include "simp_device.h"
#include "simp_kernels.h"

namespace simp {

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ int nint(double v) {
  // Fortran NINT: round half away from zero
  return (int)__builtin_round(v);
}

__device__ __forceinline__ unsigned long long ord_u64(double v) {
  unsigned long long u = (unsigned long long)__double_as_longlong(v);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__device__ __forceinline__ double ord_f64(unsigned long long u) {
  return __longlong_as_double(
      (long long)((u >> 63) ? (u & 0x7fffffffffffffffull) : ~u));
}

// x**k as the reference build expands it: x*x*...*x, left to right.
template <int K, class T>
__device__ __forceinline__ T ipow(T x) {
  T r = x;
#pragma unroll
  for (int i = 1; i < K; ++i) r = r * x;
  return r;
}

constexpr double kC6 = 78.0 / (7.0 * kPi);   // 78. / (7. * pi)
constexpr float kC6f = (float)(78.0 / (7.0 * kPi));

// particle handled by this thread.  The block table is laid out so that
// the blocks the hardware sends to one XCD (block ids congruent mod 8) cover
// a contiguous run of particles, which keeps their neighbours in that XCD's
// L2.  Workgroups go to the XCDs round-robin, continuing from where the
// previous dispatch stopped, so every grid is a multiple of 8 blocks: then a
// block id lands on the same XCD in every kernel, and the data one kernel
// writes for a run of particles is in the L2 the next kernel reads it from.
// The padding blocks (id >= the real count) exit at once.
__device__ __forceinline__ bool particle_of(const DevState &d, int &p, int &k) {
  if (blockIdx.x >= (unsigned)d.nblk) return false;
  const BlockEntry e = d.blk[blockIdx.x];
  k = e.inst;
  if ((int)threadIdx.x >= e.nloc) return false;
  p = e.pos + (int)threadIdx.x;
  return true;
}

// An instance's state is in set sel at the start of an evaluation and in
// set sel ^ rb once its lists are rebuilt: buffer b of set cs is array
// 2 cs + b of x ... hsml; mass, orig and cellid are by set.
__device__ __forceinline__ int cur_set(const InstDyn &y) { return y.sel ^ y.rb; }

// Per wavefront, the lanes holding equal keys in consecutive lanes form
// runs; one atomic add of the run's total per run.  Returns each lane's
// offset: the counter's old value plus the increments of the run's
// earlier lanes (lanes with key < 0 are not counted).
__device__ __forceinline__ int run_add(int *base, int key, int inc) {
  const int lane = (int)threadIdx.x & 63;
  const int prev = __shfl_up(key, 1, 64);
  const bool head = lane == 0 || prev != key;
  const unsigned long long heads = __ballot(head);
  // first lane of this lane's run, first lane of the next run
  const unsigned long long below = heads & ((2ull << lane) - 1ull);
  const int start = 63 - __clzll(below);
  const unsigned long long above = heads & ~((2ull << lane) - 1ull);
  const int next = above ? __ffsll((long long)above) - 1 : 64;
  // increments before this lane in its run (inclusive scan within runs)
  int incl = inc;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(incl, o, 64);
    if (lane - o >= start) incl += t;
  }
  const int total = __shfl(incl, next - 1, 64);
  int old = 0;
  if (head && key >= 0 && total != 0) old = atomicAdd(base + key, total);
  old = __shfl(old, start, 64);
  return old + incl - inc;
}

// Wave-wide min / max of a double (every lane active): DPP row shifts by
// 1, 2, 4, 8 leave each row's extreme in its lane 15, the row broadcasts
// of lanes 15 and 31 carry them into lane 63, which is read back to every
// lane.  A lane with no source keeps its own value (min and max are
// idempotent).  No LDS round trips, unlike a shuffle butterfly.
template <int ctrl, int row_mask = 0xf>
__device__ __forceinline__ double dpp_d(double v) {
  return __builtin_amdgcn_update_dpp(v, v, ctrl, row_mask, 0xf, false);
}
__device__ __forceinline__ double readlane63(double v) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <bool MAX>
__device__ __forceinline__ double wave_ext(double v) {
  auto op = [](double a, double b) { return MAX ? fmax(a, b) : fmin(a, b); };
  v = op(v, dpp_d<0x111>(v));        // row_shr:1
  v = op(v, dpp_d<0x112>(v));        // row_shr:2
  v = op(v, dpp_d<0x114>(v));        // row_shr:4
  v = op(v, dpp_d<0x118>(v));        // row_shr:8
  v = op(v, dpp_d<0x142, 0xa>(v));   // row_bcast:15 into rows 1 and 3
  v = op(v, dpp_d<0x143, 0xc>(v));   // row_bcast:31 into rows 2 and 3
  return readlane63(v);
}
__device__ __forceinline__ double wave_min(double v) { return wave_ext<false>(v); }
__device__ __forceinline__ double wave_max(double v) { return wave_ext<true>(v); }
__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Exclusive scan across the block (NT threads); returns the exclusive
// prefix of v and the block total in *total.
template <int NT = kTPB>
__device__ int block_exclusive_scan(int v, int *total) {
  __shared__ int wsum[NT / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int incl = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  if (lane == 63) wsum[wid] = incl;
  __syncthreads();
  int pre = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) {
    int s = wsum[w];
    if (w < wid) pre += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return pre + incl - v;
}

__device__ int block_sum(int v) {
  __shared__ int wsum[kTPB / 64];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
  __syncthreads();
  int tot = 0;
#pragma unroll
  for (int w = 0; w < kTPB / 64; ++w) tot += wsum[w];
  __syncthreads();
  return tot;
}

// An instance is skipped by every kernel once it has an error or is frozen.
__device__ __forceinline__ bool inst_off(const InstDyn &y) { return y.err != 0 || y.active == 0; }

// ------------------------------------------------------ reductions / grid

// Per-block min/max of x, y, min of hsml, and max of the speed and of the
// hsml change (the candidate lists' movement bound; invalid lanes give
// neutral values), stored at the block's logical id; with NP = kParts also
// the minima of hsml/|a| and hsml/c the Monaghan time step takes (only the
// kernels that end a step of such an instance compute them, and only the
// finalize of such an instance reads them).  Non-finite input raises the
// instance's NaN flag.
template <int NP = 7>
__device__ __forceinline__ void block_partials(const DevState &d, int k, bool ok, double xv,
                                               double yv, double hv, double sp, double dh,
                                               double fo = 0.0, double fc = 0.0) {
  static_assert(NP == 7 || NP == kParts, "partials: the grid's, or the grid's and dt's");
  __shared__ double sh[NP][kTPB / 64];
  bool bad = ok && (!(xv == xv) || !(yv == yv) || !(hv > 0.0) || __builtin_isinf(xv) ||
                    __builtin_isinf(yv) || __builtin_isinf(hv));
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(&d.dyn[k].err, kErrNan);
  const double big = __builtin_huge_val();
  double v[NP];
  v[0] = wave_min(ok ? xv : big); v[1] = wave_max(ok ? xv : -big);
  v[2] = wave_min(ok ? yv : big); v[3] = wave_max(ok ? yv : -big);
  v[4] = wave_min(ok ? hv : big);
  v[5] = wave_max(ok ? sp : 0.0); v[6] = wave_max(ok ? dh : 0.0);
  if (NP == kParts) {
    v[NP - 2] = wave_min(ok ? fo : big);
    v[NP - 1] = wave_min(ok ? fc : big);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int f = 0; f < NP; ++f) sh[f][w] = v[f];
  __syncthreads();
  if (threadIdx.x < NP) {
    const int f = threadIdx.x;
    const bool mx = f == 1 || f == 3 || f == 5 || f == 6;
    double r = sh[f][0];
    for (int i = 1; i < kTPB / 64; ++i) r = mx ? fmax(r, sh[f][i]) : fmin(r, sh[f][i]);
    d.part[(size_t)f * d.nblk + d.blk[blockIdx.x].lid] = r;
  }
}

// The terms of Monaghan_time_step (time_step.f90:79-83, :117) at one
// particle without artificial viscosity: hsml / |a| with a the evaluation's
// dvxdt, and hsml / c with c from c_yang (sound_speed.f90) at the step's
// final density.
// mass_update_from_area (mass_evolution.f90:66-69): rho_ice A h pi
// (alpha hsml)^2, alpha = 1. / (2 hsml_multiplicity) a default-REAL quotient
// (data_variables.f90:169)
__device__ __forceinline__ double area_mass(double A, double h, double hs) {
  const double t = (double)(1.0f / 6) * hs;
  return kRhoIce * A * h * (kPi * (t * t));
}
__device__ __forceinline__ double sound_speed(double rho) {
  // c_yang (sound_speed.f90:57-58)
  return sqrt(kYoung * (1 - kNu) / (rho * (1 + kNu) * (1 - 2 * kNu)));
}
__device__ __forceinline__ void monaghan_terms(double hs, double rho, double ax, double ay,
                                               double &fo, double &fc) {
  const double acc = sqrt(ax * ax + ay * ay);
  const double c = sound_speed(rho);
  fo = hs / acc;
  fc = hs / c;
}
// With artificial viscosity the CFL term (time_step.f90:105-107) takes the
// velocity divergence div_v of the last evaluation's pairs at the step's
// final velocities; 0.3 and 1.2 are default-REAL literals.
__device__ __forceinline__ double monaghan_cfl_av(double hs, double rho, double div_v) {
  const double c = sound_speed(rho);
  const double ad = fabs(div_v);
  return (double)0.3f * hs / (c + hs * ad + (double)1.2f * (kAlphaPi * c + kBetaPi * hs * ad));
}

// For one instance, by one block: the cells of the boundary particles (whose
// positions never change) under the grid of the next evaluation.  When none moved to
// another cell since the lists were built, nothing is done.  Otherwise
// every cell a boundary particle's 3x3 block covers is stamped with a new
// list stamp and gets the list of the boundary particles whose own cell lies
// in its 3x3 block, in ascending index -- the order in which the
// reference's boundary pair list reaches a particle of that cell
// (nnps.f90:451-493) -- with the slot the cell takes in each one's block.
__device__ __forceinline__ int lower_bound_key(const unsigned long long *keys, int n,
                                               unsigned long long v) {
  int a = 0, b = n;
  while (a < b) {
    const int mid = (a + b) >> 1;
    if (keys[mid] < v) a = mid + 1; else b = mid;
  }
  return a;
}

// boundary positions and keys a finalize thread loads before the grid is
// known (the first kBPre * kTPB boundary particles of the instance)
constexpr int kBPre = 4;
struct BPre {
  double x[kBPre], y[kBPre];
  int key[kBPre];
};

__device__ __forceinline__ void boundary_prefetch(const DevState &d, const InstMeta &m, BPre &q) {
#pragma unroll
  for (int u = 0; u < kBPre; ++u) {
    const int b = threadIdx.x + u * kTPB;
    const int bg = m.b_off + (b < m.nb ? b : 0);
    q.x[u] = d.xb[bg];
    q.y[u] = d.yb[bg];
    q.key[u] = d.bkey[bg];
  }
}

__device__ void boundary_lists(const DevState &d, int k, const BPre &q) {
  __shared__ unsigned long long keys[kMaxBoundaryPerInst];
  __shared__ int alloc;
  const int kBprepThreads = (int)blockDim.x;
  const InstMeta &m = d.meta[k];
  const InstDyn &y = d.dyn[k];
  const double bin = y.bin;
  const int W = y.W, xlo = y.xlo, ylo = y.ylo, stamp = y.stamp;
  int changed = 0;
  auto rekey = [&](int b, double xb, double yb, int old) {
    const int cxb = nint(xb / bin) - xlo + 1, cyb = nint(yb / bin) - ylo + 1;
    const int key = cyb * W + cxb;
    changed |= old != key;
    d.bkey[m.b_off + b] = key;
    keys[b] = ((unsigned long long)(unsigned)key << 32) | (unsigned)b;
  };
  // (k_finalize runs kTPB threads: the prefetched particles first)
#pragma unroll
  for (int u = 0; u < kBPre; ++u) {
    const int b = threadIdx.x + u * kTPB;
    if (b < m.nb) rekey(b, q.x[u], q.y[u], q.key[u]);
  }
  for (int b = threadIdx.x + kBPre * kTPB; b < m.nb; b += kTPB) {
    const int bg = m.b_off + b;
    rekey(b, d.xb[bg], d.yb[bg], d.bkey[bg]);
  }
  if (!__syncthreads_or(changed)) return;
  int P = 1;
  while (P < m.nb) P <<= 1;
  for (int i = m.nb + threadIdx.x; i < P; i += kBprepThreads) keys[i] = ~0ull;
  if (threadIdx.x == 0) alloc = 0;
  __syncthreads();
  // bitonic sort of (cell, b)
  for (int k2 = 2; k2 <= P; k2 <<= 1) {
    for (int j = k2 >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < P; i += kBprepThreads) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const unsigned long long a = keys[i], b = keys[ixj];
          const bool up = (i & k2) == 0;
          if ((a > b) == up) {
            keys[i] = b;
            keys[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  // Per covered cell, one thread (the first to stamp the cell's claim word
  // with this stamp) lists the boundary particles whose own cell lies in
  // the cell's 3x3 block -- three key ranges, one per row.
  int2 *brun = d.brun + m.c_off;
  int *blist = d.blist + 9 * (size_t)m.b_off;
  for (int idx = threadIdx.x; idx < m.nb * 9; idx += kBprepThreads) {
    const int b = idx / 9, s9 = idx - 9 * (idx / 9);
    const int c = d.bkey[m.b_off + b] + (s9 / 3 - 1) * W + (s9 % 3 - 1);
    d.bflag[m.c_off + c] = stamp;
    if (atomicExch(&d.bcellc[m.c_off + c], stamp) == stamp) continue;
    int lo[3], hi[3], total = 0;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      // cells c0 .. c0 + 2 of one row (a row outside the grid is empty;
      // the cells a row range borrows from a neighbouring row are margin
      // cells, which hold no boundary particle)
      const int c0 = c + (r - 1) * W - 1;
      lo[r] = hi[r] = 0;
      if (c0 + 3 > 0) {
        lo[r] = lower_bound_key(keys, m.nb, (unsigned long long)(unsigned)max(c0, 0) << 32);
        hi[r] = lower_bound_key(keys, m.nb, (unsigned long long)(unsigned)(c0 + 3) << 32);
      }
      total += hi[r] - lo[r];
    }
    const int start = atomicAdd(&alloc, total);
    brun[c] = make_int2(start, start + total);
    int last = -1;
    for (int j = 0; j < total; ++j) {
      int best = 0x7fffffff, bcell = 0;
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        for (int t = lo[r]; t < hi[r]; ++t) {
          const int bl = (int)(keys[t] & 0xffffffffu);
          if (bl > last && bl < best) {
            best = bl;
            bcell = (int)(keys[t] >> 32);
          }
        }
      }
      last = best;
      // slot of c in the 3x3 block of the boundary particle's cell
      const int dyc = (c - bcell + W + 1) / W;     // 0, 1, 2
      const int dxc = c - bcell - (dyc - 1) * W + 1;
      blist[start + j] = best | ((dyc * 3 + dxc) << 24);
    }
  }
  if (threadIdx.x == 0) d.dyn[k].lstamp = stamp;
}

__global__ void __launch_bounds__(kTPB) k_partial(DevState d, int buf) {
  if (blockIdx.x >= (unsigned)d.nblk) return;
  int p = 0, k = 0;
  bool ok = particle_of(d, p, k);
  const int cs = cur_set(d.dyn[k]);
  double xv = ok ? d.x[2 * cs + buf][p] : 0.0, yv = ok ? d.y[2 * cs + buf][p] : 0.0;
  double hv = ok ? d.hsml[2 * cs + buf][p] : 1.0;
  block_partials(d, k, ok, xv, yv, hv, 0.0, 0.0);
}

__device__ int finalize_grid(const InstMeta &m, InstDyn &y, int end_of_step, double xmn,
                             double xmx, double ymn, double ymx, double hmn, double mforce,
                             double mcfl);

__device__ __forceinline__ void stats_reset(InstDyn &y) {
  y.ipstep = y.itimestep;
  y.ipmax = 0;
  y.ipmin = 0x7fffffff;
  y.ipzero = 0;
  y.ipsum = 0;
  y.ipown = 0;
}

// Interaction statistics of one wave's particles (every lane takes part;
// lanes without a particle pass live = false): max, min, sum and the count
// of particles with no interaction, then the owned particle pairs.
__device__ __forceinline__ void stats_add(InstDyn &y, bool live, int cnt, int own) {
  int mx = live ? cnt : 0, mn = live ? cnt : 0x7fffffff, z = (live && cnt == 0) ? 1 : 0;
  unsigned long long sm = live ? (unsigned long long)cnt : 0ull;
  unsigned long long ow = live ? (unsigned long long)own : 0ull;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    mx = max(mx, __shfl_xor(mx, o, 64));
    mn = min(mn, __shfl_xor(mn, o, 64));
    z += __shfl_xor(z, o, 64);
    sm += __shfl_xor(sm, o, 64);
    ow += __shfl_xor(ow, o, 64);
  }
  if ((threadIdx.x & 63) == 0 && __any(live)) {
    atomicMax(&y.ipmax, mx);
    atomicMin(&y.ipmin, mn);
    atomicAdd(&y.ipzero, z);
    atomicAdd(&y.ipsum, sm);
    atomicAdd(&y.ipown, ow);
  }
}


// end of an evaluation: the current set becomes the instance's set
__device__ __forceinline__ void fold_set(InstDyn &y) {
  y.sel ^= y.rb;
  y.rb = 0;
}

// end of an evaluation of a step: lists pass A built in it are valid from
// now on (a finalize that follows no pass A -- the first grid of a new
// layout -- changes nothing)
__device__ __forceinline__ void lists_built(InstDyn &y) {
  if (y.vskin_next > 0.f) {
    y.vskin = y.vskin_next;
    y.vskin_next = 0.f;
    y.vmoved = 0.f;          // built at the state just evaluated
    y.vbuilt = 1;
  }
}

// The next state moved from the evaluated one by at most 1.5 dt times the
// largest speed of this or the previous evaluation's state (predictor: dt/2
// v_S; corrector: dt v_T - dt/2 v_S), hsml by at most the largest change;
// 2 |dr| + |dh| adds to the lists' movement bound (rounded up).
__device__ __forceinline__ void move_bound(InstDyn &y, double vmax, double dhmax) {
  const double v = fmax(vmax, (double)y.vprev);
  y.vmoved = (float)((double)y.vmoved + 3.0 * y.dt * v + dhmax) * 1.00001f;
  y.vprev = (float)vmax * 1.00001f;
}

// a single evaluation collects the interaction statistics whatever its step
__global__ void __launch_bounds__(kTPB) k_stats_reset(DevState d) {
  const int k = blockIdx.x * kTPB + threadIdx.x;
  if (k < d.ninst) stats_reset(d.dyn[k]);
}

__global__ void __launch_bounds__(kTPB) k_fold(DevState d) {
  const int k = blockIdx.x * kTPB + threadIdx.x;
  if (k < d.ninst) {
    fold_set(d.dyn[k]);
    d.dyn[k].vbuilt = 0;    // no movement measured for the next state
  }
}

// One block per instance: fold the instance's block partials into the grid
// of the next evaluation (bucket_search, nnps.f90:211-224).  With
// end_of_step it first closes the step: VP dt from MINVAL of the new hsml
// (time_step.f90:155), itimestep + 1 and current_time + dt
// (time_integration.f90:68-71).
__global__ void __launch_bounds__(kTPB) k_finalize(DevState d, int end_of_step) {
  __shared__ double sh[kParts][kTPB / 64];
  if (blockIdx.x >= (unsigned)d.ninst) return;
  const int k = blockIdx.x;
  const InstMeta &m = d.meta[k];
  // thread 0 works on a copy of the instance record, loaded with the
  // partials (one round of loads, not one dependent round trip per field
  // it reads) and stored back whole once the grid is set
  InstDyn y;
  if (threadIdx.x == 0) y = d.dyn[k];

  // the boundary particles' positions and keys go out first: the boundary
  // lists below need them once the grid is known
  BPre bq;
  if (m.nb > 0) boundary_prefetch(d, m, bq);
  const double big = __builtin_huge_val();
  double v0 = big, v1 = -big, v2 = big, v3 = -big, v4 = big, v5 = 0.0, v6 = 0.0;
  double v7 = big, v8 = big;
  const size_t nb = (size_t)d.nblk;
  // the Monaghan time step's minima exist at the end of such a step only
  const bool mdt = end_of_step && m.tse == 1;
  // kFin block partials per thread and round: their loads go out together
  constexpr int kFin = 4;
  const int bend = m.blk_off + m.nblk;
  for (int b0 = m.blk_off + threadIdx.x; b0 < bend; b0 += kFin * kTPB) {
    double a[kFin][7];
#pragma unroll
    for (int u = 0; u < kFin; ++u) {
      const int b = b0 + u * kTPB < bend ? b0 + u * kTPB : b0;
#pragma unroll
      for (int j = 0; j < 7; ++j) a[u][j] = d.part[j * nb + b];
    }
#pragma unroll
    for (int u = 0; u < kFin; ++u) {
      // (a repeated partial leaves a min / max unchanged)
      v0 = fmin(v0, a[u][0]);
      v1 = fmax(v1, a[u][1]);
      v2 = fmin(v2, a[u][2]);
      v3 = fmax(v3, a[u][3]);
      v4 = fmin(v4, a[u][4]);
      v5 = fmax(v5, a[u][5]);
      v6 = fmax(v6, a[u][6]);
    }
    if (mdt) {
      for (int b = b0; b < bend && b < b0 + kFin * kTPB; b += kTPB) {
        v7 = fmin(v7, d.part[7 * nb + b]);
        v8 = fmin(v8, d.part[8 * nb + b]);
      }
    }
  }
  v0 = wave_min(v0); v1 = wave_max(v1); v2 = wave_min(v2); v3 = wave_max(v3);
  v4 = wave_min(v4); v5 = wave_max(v5); v6 = wave_max(v6);
  v7 = wave_min(v7); v8 = wave_min(v8);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sh[0][w] = v0; sh[1][w] = v1; sh[2][w] = v2; sh[3][w] = v3; sh[4][w] = v4;
    sh[5][w] = v5; sh[6][w] = v6; sh[7][w] = v7; sh[8][w] = v8;
  }
  __syncthreads();
  __shared__ int go;
  if (threadIdx.x == 0) {
    for (int i = 1; i < kTPB / 64; ++i) {
      v0 = fmin(sh[0][0], sh[0][i]); sh[0][0] = v0;
      v1 = fmax(sh[1][0], sh[1][i]); sh[1][0] = v1;
      v2 = fmin(sh[2][0], sh[2][i]); sh[2][0] = v2;
      v3 = fmax(sh[3][0], sh[3][i]); sh[3][0] = v3;
      v4 = fmin(sh[4][0], sh[4][i]); sh[4][0] = v4;
      v5 = fmax(sh[5][0], sh[5][i]); sh[5][0] = v5;
      v6 = fmax(sh[6][0], sh[6][i]); sh[6][0] = v6;
      v7 = fmin(sh[7][0], sh[7][i]); sh[7][0] = v7;
      v8 = fmin(sh[8][0], sh[8][i]); sh[8][0] = v8;
    }
    fold_set(y);
    if (!inst_off(y)) lists_built(y);
    move_bound(y, sh[5][0], sh[6][0]);
    go = finalize_grid(m, y, end_of_step, sh[0][0], sh[1][0], sh[2][0], sh[3][0], sh[4][0],
                       sh[7][0], sh[8][0]);
    d.dyn[k] = y;
  }
  __syncthreads();
  // the boundary particles' cell lists for the next evaluation
  if (go && m.nb > 0) boundary_lists(d, k, bq);
}

// Thread 0 of k_finalize: close the step, then the grid of the next
// evaluation.  Returns whether the instance evaluates next.
__device__ int finalize_grid(const InstMeta &m, InstDyn &y, int end_of_step, double xmn,
                             double xmx, double ymn, double ymx, double hmn, double mforce,
                             double mcfl) {
  if (inst_off(y)) return 0;
  if (end_of_step) {
    // VP_time_step: MINVAL(rho_ice hsml^2 delta_min / (P_star (1 + k_t))) is
    // the same expression at MINVAL(hsml), each operation being monotone
    if (m.tse == 2) y.dt = kRhoIce * (hmn * hmn) * kDeltaMin / (kPStar * (1 + kKT));
    // Monaghan_time_step (time_step.f90:79-121): fF = 0.25 times the root of
    // MINVAL(hsml / |a|), or MINVAL(hsml / c) if smaller
    if (m.tse == 1) {
      const double dt_force = 0.25 * sqrt(mforce);
      y.dt = dt_force < mcfl ? dt_force : mcfl;
    }
    y.itimestep += 1;
    y.time = y.time + y.dt;
    if (y.steps_left > 0 && --y.steps_left == 0) {
      y.active = 0;   // finished: frozen from the next evaluation on
      return 0;
    }
  }
  if (m.nb > 0) {
    xmn = fmin(xmn, m.bxmin); xmx = fmax(xmx, m.bxmax);
    ymn = fmin(ymn, m.bymin); ymx = fmax(ymx, m.bymax);
  }
  const double bin = 2 * hmn;
  const int xlo = nint(xmn / bin), xhi = nint(xmx / bin);
  const int ylo = nint(ymn / bin), yhi = nint(ymx / bin);
  const long long W = (long long)xhi - xlo + 3, H = (long long)yhi - ylo + 3;
  if (!(hmn > 0.0) || W * H > (long long)m.c_cap || W * H <= 0) {
    y.err |= kErrCapacity;
    return 0;
  }
  y.hmin = hmn;
  y.bin = bin;
  // FP32 coordinates relative to the origin are within (max(W, H) + 1) bin;
  // each rounds by at most 2^-24 of that, a difference by twice as much, and
  // mhsml by 2^-24 relative: 8 ulp of the extent bounds the distance error
  y.pf_slack = (float)(8.0 * 5.9604644775390625e-8 * (double)(max(W, H) + 1) * bin + 1e-3);
  y.xlo = xlo;
  y.ylo = ylo;
  y.W = (int)W;
  y.H = (int)H;
  y.stamp += 1;
  // the evaluations of a print step collect the interaction statistics
  if (y.itimestep % kPrintStep == 0) stats_reset(y);
  return 1;
}

// ------------------------------------------------------------------ count

// The inputs of pass A's prefilter for slot p (lanes with want; every lane
// of the wavefront takes part): its FP32 position relative to the grid
// origin and hsml, and per cell the largest member smoothing length
// (rounded up to FP32), the bound the prefilter's cell culling uses: a max
// over each run of equal keys in the wavefront (slots are in cell order,
// but for particles that just changed cell), then one 64-bit atomic max
// per run of (evaluation stamp, value) -- a stale stamp reads as an empty
// cell, so nothing needs clearing.
__device__ __forceinline__ void prefilter_inputs(const DevState &d, bool want, int k, int p,
                                                 int key, double xq, double yq, double hq) {
  const int lane = (int)threadIdx.x & 63;
  int kk = -1 - lane, coff = 0;
  unsigned long long hv = 0;
  if (want) {
    const InstDyn &y = d.dyn[k];
    const double ox = (y.xlo - 1) * y.bin, oy = (y.ylo - 1) * y.bin;
    d.pf[p] = make_float4((float)(xq - ox), (float)(yq - oy), (float)hq, 0.0f);
    kk = key;
    coff = d.meta[k].c_off;
    hv = ((unsigned long long)(unsigned)y.stamp << 32) | __float_as_uint((float)hq * 1.000001f);
  }
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned lo = __shfl_down((unsigned)hv, o, 64);
    const unsigned hi = __shfl_down((unsigned)(hv >> 32), o, 64);
    const int ok = __shfl_down(kk, o, 64);
    const unsigned long long ov = ((unsigned long long)hi << 32) | lo;
    if (lane + o < 64 && ok == kk && ov > hv) hv = ov;
  }
  const int prev = __shfl_up(kk, 1, 64);
  if (want && (lane == 0 || prev != kk)) atomicMax(&d.cellh[coff + kk], hv);
}

// pass A runs its prefilter in this evaluation unless the cell lists are
// rebuilt (then the state move supplies the prefilter's inputs)
__device__ __forceinline__ bool lists_stale(const InstDyn &y) {
  return !y.vbuilt || !(y.vmoved < 0.9f * y.vskin);
}

// Cell key of every slot (bucket_search, nnps.f90:249-250), and the
// prefilter's inputs when pass A will prefilter.
// The member counts of the cells follow the keys that change.  An
// instance's lists are rebuilt only when one of its keys differs from the
// one its slot was sorted with (or on its first evaluation); the rebuild
// kernels below return at once for the others.  Between the predictor and
// the corrector evaluation of a step the grid and the keys normally stay
// the same, so the corrector reuses the lists.
__global__ void __launch_bounds__(kTPB) k_count(DevState d, int buf) {
  int p = 0, k = 0;
  bool changed = false;
  const int lane = (int)threadIdx.x & 63;
  int key = -1 - lane, oldk = -1 - lane;
  const bool valid = particle_of(d, p, k) && !inst_off(d.dyn[k]);
  int coff = 0;
  double xq = 0., yq = 0., hq = 0.;
  if (valid) {
    const InstDyn &y = d.dyn[k];
    const int cs = y.sel;
    coff = d.meta[k].c_off;
    const double bin = y.bin;
    xq = d.x[2 * cs + buf][p];
    yq = d.y[2 * cs + buf][p];
    hq = d.hsml[2 * cs + buf][p];
    const int cx = nint(xq / bin) - y.xlo + 1;
    const int cy = nint(yq / bin) - y.ylo + 1;
    const int c = cy * y.W + cx;
    const int old = d.cellid[cs][p];
    changed = old != c;
    if (changed) {
      d.cellid[cs][p] = c;
      oldk = old;   // < 0 on the first evaluation
    }
    key = c;
  }
  // blocks never straddle instances: one instance per wavefront.  With
  // force_search every evaluation rebuilds the lists (the same order: the
  // measurement of the reference's search at every evaluation).
  // (plain stores of the same value: atomics from every wavefront of a
  // rebuild would serialize on two addresses)
  if ((d.force_search ? __any(valid) : __any(changed)) && lane == 0) {
    d.dyn[k].rb = 1;
    *d.rebuild = 1;
  }
  if (__any(changed)) {
    run_add(d.count + coff, changed ? key : -1 - lane, 1);
    run_add(d.count + coff, oldk, -1);
  }
  prefilter_inputs(d, valid && lists_stale(d.dyn[k]), k, p, key, xq, yq, hq);
}

// ------------------------------------------------------------------- scan

__global__ void __launch_bounds__(kTPB) k_scan1(const int *in, int L, int *bsum, int nblocks,
                                                const int *gate) {
  if (blockIdx.x >= (unsigned)nblocks || (gate && *gate == 0)) return;
  constexpr int IPT = kScanItems / kTPB;
  int base = blockIdx.x * kScanItems + threadIdx.x * IPT;
  int s = 0;
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    int idx = base + i;
    if (idx < L) s += in[idx];
  }
  s = block_sum(s);
  if (threadIdx.x == 0) bsum[blockIdx.x] = s;
}

// out[i] = sum(in[0..i-1]); out[L] = total.  Optionally clears in[].
__global__ void __launch_bounds__(kTPB) k_scan2(int *in, int L, const int *bsum, int *out,
                                                int zero_in, int nblocks, const int *gate) {
  if (blockIdx.x >= (unsigned)nblocks || (gate && *gate == 0)) return;
  constexpr int IPT = kScanItems / kTPB;
  int acc = 0;
  for (int j = threadIdx.x; j < (int)blockIdx.x; j += kTPB) acc += bsum[j];
  int pre = block_sum(acc);
  int base = blockIdx.x * kScanItems + threadIdx.x * IPT;
  int v[IPT];
  int s = 0;
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    int idx = base + i;
    v[i] = idx < L ? in[idx] : 0;
    s += v[i];
  }
  int tot = 0;
  int ex = block_exclusive_scan(s, &tot) + pre;
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    int idx = base + i;
    if (idx < L) {
      out[idx] = ex;
      if (zero_in) in[idx] = 0;
    }
    ex += v[i];
  }
  if (blockIdx.x == (unsigned)nblocks - 1 && threadIdx.x == kTPB - 1) out[L] = ex;
}

// --------------------------------------------------------------- rebuild
//
// The rebuild of the cell lists is one kernel of at most kRebuildBlocks
// blocks (at most four per CU: every block of it is resident at once) whose
// blocks stride over the block table.  Between rebuilds -- most
// evaluations -- it reads one flag and returns: one launch instead of the
// four a rebuild needs.  A rebuild runs four phases separated by grid
// barriers: (1) slots in the new cells and the block sums of the cell
// counts, (2) the cell starts, (3) the new order, (4) the state move.
//
// The data crossing a barrier (block sums, slots, cell starts, the new
// order) is written and read with agent-scope atomic stores and loads,
// which reach memory past the XCDs' L2 caches; a barrier waits until each
// block's stores have completed before the block arrives.  No L2 writeback
// or invalidate is needed (an agent-scope fence per barrier costs ~30 us
// on MI355X, this barrier a few).

constexpr int kRebuildBlocks = 1024;   // at most; four per CU on MI355X
constexpr int kBarSpin = 1 << 18;   // bound of a barrier's wait (polls)
// barrier words: kBarGroups arrival counters (one 128-byte line each), then
// the counter of the groups arrived and the barrier generation
constexpr int kBarGroups = 8;
constexpr int kBarStride = 32;
static_assert(kBarWords == kBarGroups * kBarStride + 2, "barrier words");

__device__ __forceinline__ int ld_ag(const int *p) {
  return __hip_atomic_load(const_cast<int *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_ag(int *p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid barrier in two levels: the blocks of group blockIdx % kBarGroups
// arrive on their group's counter, the last of each group on the top
// counter, and the last group advances the generation (every counter is
// back to zero by then).  Returns false when the wait ran past its bound
// (the kernel then flags every instance).
__device__ bool grid_sync(unsigned *bar) {
  __shared__ int ok;
  __builtin_amdgcn_s_waitcnt(0);     // this thread's stores have completed
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned grp = blockIdx.x % kBarGroups;
    const unsigned gsize = (gridDim.x + kBarGroups - 1 - grp) / kBarGroups;
    unsigned *count = bar + grp * kBarStride;
    unsigned *top = bar + kBarGroups * kBarStride, *gen = top + 1;
    const unsigned g = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);   // gen read before arriving
    bool last = false;
    if (__hip_atomic_fetch_add(count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
        gsize - 1) {
      __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __builtin_amdgcn_s_waitcnt(0);
      const unsigned ng = min((unsigned)kBarGroups, gridDim.x);
      last = __hip_atomic_fetch_add(top, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
             ng - 1;
    }
    int good = 1;
    if (last) {
      __hip_atomic_store(top, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __builtin_amdgcn_s_waitcnt(0);
      __hip_atomic_fetch_add(gen, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      int it = 0;
      while (__hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
        if (++it > kBarSpin) {
          good = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    ok = good;
  }
  __syncthreads();
  return ok != 0;
}

__device__ __forceinline__ bool particle_at(const DevState &d, int vb, int &p, int &k) {
  const BlockEntry e = d.blk[vb];
  k = e.inst;
  if ((int)threadIdx.x >= e.nloc) return false;
  p = e.pos + (int)threadIdx.x;
  return true;
}

// The member of cell c (listed in cl's range [s0, s0 + n) in arbitrary
// order) whose particle id ranks r among the cell's members: the cell's
// members in ascending particle id are the order in which the reference
// fills its bins (nnps.f90:246-254).  Cells of up to kRankMax members (the
// common case) are ranked in registers.
constexpr int kRankMax = 16;

__device__ __forceinline__ int cell_member_at(const int *cl, const int *orig, int s0, int n,
                                              int r) {
  if (n == 1) return ld_ag(cl + s0);
  if (n <= kRankMax) {
    int v[kRankMax], key[kRankMax];
#pragma unroll
    for (int i = 0; i < kRankMax; ++i) v[i] = i < n ? ld_ag(cl + s0 + i) : 0;
#pragma unroll
    for (int i = 0; i < kRankMax; ++i) key[i] = i < n ? orig[v[i]] : 0x7fffffff;
    int q = v[0];
#pragma unroll
    for (int i = 0; i < kRankMax; ++i) {
      int rk = 0;
#pragma unroll
      for (int j = 0; j < kRankMax; ++j) rk += key[j] < key[i];
      if (i < n && rk == r) q = v[i];
    }
    return q;
  }
  for (int i = 0; i < n; ++i) {
    const int vi = ld_ag(cl + s0 + i), ki = orig[vi];
    int rk = 0;
    for (int j = 0; j < n; ++j) rk += orig[ld_ag(cl + s0 + j)] < ki;
    if (rk == r) return vi;
  }
  return ld_ag(cl + s0);
}

// bsum: nsb block sums of the cell counts (kScanItems cells each)
__global__ void __launch_bounds__(kTPB) k_rebuild(DevState d, int *bsum, int nsb, int buf) {
  if (*d.rebuild == 0) return;       // the same for every block: no rebuild
  const int G = (int)gridDim.x;
  constexpr int IPT = kScanItems / kTPB;
  const int L = d.ctot;
  bool ok = true;

  // (1) each particle's slot in its new cell (consecutive slots mostly share
  //     a cell: one atomic per run; fill is zero again after the state
  //     move), and the block sums of the cell counts
  for (int vb = blockIdx.x; vb < max(nsb, d.nblk); vb += G) {
    if (vb < nsb) {
      const int base = vb * kScanItems + threadIdx.x * IPT;
      int sum = 0;
#pragma unroll
      for (int i = 0; i < IPT; ++i)
        if (base + i < L) sum += d.count[base + i];
      sum = block_sum(sum);
      if (threadIdx.x == 0) st_ag(bsum + vb, sum);
    }
    if (vb < d.nblk && d.dyn[d.blk[vb].inst].rb != 0) {   // one instance per block
      int p = 0, k = 0;
      const int lane = (int)threadIdx.x & 63;
      int key = -1 - lane, coff = 0;
      if (particle_at(d, vb, p, k)) {
        coff = d.meta[k].c_off;
        key = d.cellid[d.dyn[k].sel][p];
      }
      const int sl = run_add(d.fill + coff, key, 1);
      if (key >= 0) st_ag(d.slot + p, sl);
    }
  }
  ok = grid_sync(d.bar) && ok;

  // (2) cell starts: cell_start[i] = counts before cell i, cell_start[L]
  //     the total
  for (int vb = blockIdx.x; vb < nsb; vb += G) {
    int acc = 0;
    for (int j = threadIdx.x; j < vb; j += kTPB) acc += ld_ag(bsum + j);
    const int pre = block_sum(acc);
    const int base = vb * kScanItems + threadIdx.x * IPT;
    int v[IPT];
    int s = 0;
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
      v[i] = base + i < L ? d.count[base + i] : 0;
      s += v[i];
    }
    int tot = 0;
    int ex = block_exclusive_scan(s, &tot) + pre;
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
      if (base + i < L) st_ag(d.cell_start + base + i, ex);
      ex += v[i];
    }
    if (vb == nsb - 1 && threadIdx.x == kTPB - 1) st_ag(d.cell_start + L, ex);
  }
  ok = grid_sync(d.bar) && ok;

  // (3) the new order: cl[new slot] = old slot
  for (int vb = blockIdx.x; vb < d.nblk; vb += G) {
    int p = 0, k = 0;
    if (!particle_at(d, vb, p, k)) continue;
    const InstDyn &y = d.dyn[k];
    if (y.rb == 0) continue;
    const int c = d.meta[k].c_off + d.cellid[y.sel][p];
    st_ag(d.cl + ld_ag(d.cell_start + c) + ld_ag(d.slot + p), p);
  }
  ok = grid_sync(d.bar) && ok;
  if (!ok) {
    // a barrier's wait ran out: the rebuilt lists cannot be trusted
    for (int k = blockIdx.x * kTPB + threadIdx.x; k < d.ninst; k += G * kTPB)
      atomicOr(&d.dyn[k].err, kErrSync);
    return;
  }

  // (4) the state of a rebuilt instance moves to the new slot order, from
  //     set sel to set sel ^ 1: each slot takes the member of its cell that
  //     ranks there by particle id (its old slot), and writes the prefilter
  //     positions of the evaluation (buffer buf).  The half-step buffer
  //     moves only when it is live (the corrector's evaluation).  The first
  //     slot of each cell clears the cell's fill counter for the next
  //     rebuild.
  for (int vb = blockIdx.x; vb < d.nblk; vb += G) {
    if (d.dyn[d.blk[vb].inst].rb == 0) continue;   // one instance per block
    if (threadIdx.x == 0) d.blk[vb].cs = d.dyn[d.blk[vb].inst].sel ^ 1;
    int s = 0, k = 0;
    const bool act = particle_at(d, vb, s, k);
    int key = 0;
    double xq = 0., yq = 0., hq = 0.;
    if (act) {
      const InstDyn &y = d.dyn[k];
      const int f = y.sel, t = y.sel ^ 1;
      const int coff = d.meta[k].c_off;
      // the cell of slot s: that of any particle placed in its range
      const int c = d.cellid[f][ld_ag(d.cl + s)];
      const int s0 = ld_ag(d.cell_start + coff + c), s1 = ld_ag(d.cell_start + coff + c + 1);
      if (s == s0) d.fill[coff + c] = 0;
      const int q = cell_member_at(d.cl, d.orig[f], s0, s1 - s0, s - s0);
      for (int b = 0; b <= buf; ++b) {
        const int i = 2 * f + b, o = 2 * t + b;
        d.x[o][s] = d.x[i][q]; d.y[o][s] = d.y[i][q];
        d.vx[o][s] = d.vx[i][q]; d.vy[o][s] = d.vy[i][q];
        d.h[o][s] = d.h[i][q]; d.A[o][s] = d.A[i][q];
        d.rho[o][s] = d.rho[i][q]; d.hsml[o][s] = d.hsml[i][q];
      }
      d.mass[t][s] = d.mass[f][q];
      if (buf == 1 && d.masst[0]) d.masst[t][s] = d.masst[f][q];
      d.orig[t][s] = d.orig[f][q];
      key = c;
      d.cellid[t][s] = key;
      const int i = 2 * f + buf;
      xq = d.x[i][q];
      yq = d.y[i][q];
      hq = d.hsml[i][q];
    }
    // pass A prefilters after a rebuild: its inputs in the new order
    prefilter_inputs(d, act, k, s, key, xq, yq, hq);
  }
}

// -------------------------------------------------------------- boundary

// --------------------------------------------------------------- neighbours

// The 3x3 cell block of a particle as six contiguous ranges of the cell
// lists, in the reference's pair order for that particle.  kind: 0 = pair
// owned by the neighbour's cell (or by an earlier member of the own cell),
// 1 = own-cell pair owned by this particle (intra, nnps.f90:265-297),
// 2 = cross-cell pair owned by this particle (right, up, up-right, up-left;
// nnps.f90:300-407).
struct Ranges {
  int a[6], b[6];
};

__device__ __forceinline__ Ranges neighbour_ranges(const int *cs, int W, int cx, int cy) {
  Ranges r;
  const int rm = (cy - 1) * W, r0 = cy * W, rp = (cy + 1) * W;
  r.a[0] = cs[rm + cx - 1]; r.b[0] = cs[rm + cx + 2];   // row below
  r.a[1] = cs[r0 + cx - 1]; r.b[1] = cs[r0 + cx];       // left
  r.a[2] = r.b[1];          r.b[2] = cs[r0 + cx + 1];   // own cell
  r.a[3] = r.b[2];          r.b[3] = cs[r0 + cx + 2];   // right
  r.a[4] = cs[rp + cx];     r.b[4] = cs[rp + cx + 2];   // up, up-right
  r.a[5] = cs[rp + cx - 1]; r.b[5] = r.a[4];            // up-left
  return r;
}

// The six ranges form four contiguous segments of the cell lists (the left,
// own and right cells of a row are adjacent): [a0,b0) [a1,b3) [a4,b4)
// [a5,a4), walked in that order.  Rows lie in the lists in ascending order,
// so a list position alone tells its kind: below the middle row 0, in the
// top row 2, in the middle row 0 / 1 / 2 by left-or-earlier / later own-cell
// member / right.
// neighbour loads issued together, the defaults of the tuning parameters
// (pass A's exact test; pass B FP64 and mixed)
constexpr int kBatchA = 2;
constexpr int kBatchBFP64 = 4;
constexpr int kBatchBMixed = 3;

__device__ __forceinline__ int pair_kind(const Ranges &R, int s, int self) {
  if (s < R.a[1]) return 0;
  if (s >= R.a[5]) return 2;
  if (s < R.a[2]) return 0;
  if (s < R.b[2]) return s > self ? 1 : 0;
  return 2;
}

// Exact FP64 form of `d = SQRT(SUM(dx**2)); IF (d .LE. mhsml)`: the square
// root is taken only when d^2 is within 1e-13 of mhsml^2, which decides the
// far candidates without it.
__device__ __forceinline__ bool in_support(double d2, double mh, double &r) {
  if (d2 > (mh * mh) * (1.0 + 1e-13)) return false;
  r = sqrt(d2);
  return r <= mh;
}

// The same predicate without d: below (1 - 1e-13) mhsml^2 the rounded
// square root cannot exceed mhsml, so only the band between takes it.
__device__ __forceinline__ bool in_support_nr(double d2, double mh) {
  const double m2 = mh * mh;
  if (d2 > m2 * (1.0 + 1e-13)) return false;
  if (d2 < m2 * (1.0 - 1e-13)) return true;
  return sqrt(d2) <= mh;
}

// Wendland C6 gradient factor (kernel.f90:499-508):
// dW/dx = g * dx / r with g = factor * (-22 q (1-q)^7 (16 q^2 + 7 q + 1)) / hsml_ij
__device__ __forceinline__ double c6_grad(double r, double mh) {
  double q = r / mh * 1.0;
  double ih = 1.0 / mh;
  double factor = kC6 * (ih * ih);
  double omq = 1 - q;
  double o7 = ipow<7>(omq);
  return factor * (-(22. * q * o7 * (16 * ipow<2>(q) + 7. * q + 1))) / mh;
}

// Wendland C6 kernel value and the factor kernel_zero stores
// (kernel.f90:499-506)
__device__ __forceinline__ double c6_factor(double mh) {
  double ih = 1.0 / mh;
  return kC6 * (ih * ih);
}

__device__ __forceinline__ double c6_w(double r, double mh) {
  double q = r / mh * 1.0;
  double ih = 1.0 / mh;
  double factor = kC6 * (ih * ih);
  double omq = 1 - q;
  return factor * ipow<8>(omq) * (32 * ipow<3>(q) + 25 * ipow<2>(q) + 8 * q + 1);
}

// FP32 kernel gradient over the distance, from the squared distance:
// grad W = (dx, dy) * c6_gradr_f(dx^2 + dy^2, mh), the same polynomial
// with the 1/r folded in (hardware sqrt and reciprocal, ~1 ulp)
__device__ __forceinline__ float c6_gradr_f(float d2, float mh) {
  const float ih = __builtin_amdgcn_rcpf(mh);
  const float q = __builtin_amdgcn_sqrtf(d2) * ih;
  const float ih2 = ih * ih;
  const float omq = 1.0f - q;
  const float o2 = omq * omq, o4 = o2 * o2;
  const float o7 = o4 * o2 * omq;
  const float poly = __builtin_fmaf(__builtin_fmaf(16.0f, q, 7.0f), q, 1.0f);
  return (-22.0f * kC6f) * (ih2 * ih2) * o7 * poly;
}

// The Wendland C2 and C4 kernels (skf = 4, 5; kernel.f90:387-396,
// :443-452; scale_h = 1) beside C6, in the reference's evaluation order:
// the gradient factor g (dW/dx = g dx / r), the value w, kernel_zero (the
// value at r = 0) and the FP32 gradient over the distance.
constexpr double kC2 = 3.0 / (2.0 * kPi);    // 3. / (2. * pi)
constexpr double kC4 = 3.0 / kPi;            // 3. / pi
constexpr float kC2f = (float)(3.0 / (2.0 * kPi));
constexpr float kC4f = (float)(3.0 / kPi);

// scale_h of each kernel (initialize.f90:207-217): q = scale_h r / h and
// dW/dx = g (dx / r) scale_h; the support stays mhsml
__device__ __forceinline__ double kern_scale(int skf) {
  return skf == 1 ? 2.0 : (skf == 2 || skf == 3 || skf == 7 || skf == 8) ? 3.0 : 1.0;
}

// The quadratic (skf = 1, kernel.f90:269-278), Gaussian (2, :136-145),
// quintic spline (3, :193-221) and Marquis kernels (7, :555-564; 8,
// :604-621): value w, gradient factor g and kernel_zero of one pair, FP64
// in both precision modes.  Marquis2's 34.1558 and log(2.) are default-REAL.
constexpr double kLn2f = (double)0x1.62e43p-1f;       // log(2.) in REAL(4)
constexpr double kMarquis2 = (double)0x1.113f14p+5f;  // 34.1558 in REAL(4)

__device__ void kern_other(int skf, double r, double mh, double *w, double *g, double *kz) {
  const double sh = kern_scale(skf);
  const double q = r / mh * sh;
  const double t = sh / mh;
  if (skf == 1) {
    const double factor = 2. / kPi * (t * t);
    if (kz) *kz = factor * 3. / 4.;
    if (w) *w = factor * (3. / 16. * ipow<2>(q) - 3. / 4. * q + 3. / 4.);
    if (g) *g = factor * (3. / 8. * q - 3. / 4.) / mh;
  } else if (skf == 2) {
    const double factor = 1. / kPi * (t * t);
    const double e = exp(-ipow<2>(q));
    if (kz) *kz = factor;
    if (w) *w = factor * e;
    if (g) *g = factor * (-(2 * q * e)) / mh;
  } else if (skf == 3) {
    const double factor = 7.e0 / (478.e0 * kPi) * (t * t);
    if (kz) *kz = factor * 66;
    const double a = 3 - q, b = 2 - q, c = 1 - q;
    double ww, gg;
    if (q >= 0. && q <= 1.) {
      ww = factor * (ipow<5>(a) - 6 * ipow<5>(b) + 15 * ipow<5>(c));
      gg = factor * (-(5 * ipow<4>(a)) + 30 * ipow<4>(b) - 75 * ipow<4>(c)) / mh;
    } else if (q > 1. && q <= 2.) {
      ww = factor * (ipow<5>(a) - 6 * ipow<5>(b));
      gg = factor * (-(5 * ipow<4>(a)) + 30 * ipow<4>(b)) / mh;
    } else {
      ww = factor * ipow<5>(a);
      gg = factor * (-(5 * ipow<4>(a))) / mh;
    }
    if (w) *w = ww;
    if (g) *g = gg;
  } else if (skf == 7) {
    const double factor = 1. / kPi * (t * t);
    if (kz) *kz = factor * 0.5;                       // (0.5 - 0.5 * tanh(0.))
    if (w) *w = factor * (0.5 - 0.5 * tanh(300 * (q - 1)));
    if (g) {
      const double ch = cosh(300 * (-1 + q));
      *g = factor * (-(150 * (1.0 / (ch * ch)))) / mh;
    }
  } else {
    const double C1 = 10., C2 = 50.;
    const double factor = 1. / (kMarquis2 * kPi * ipow<2>(mh));
    if (kz) *kz = factor * C2 * (log(cosh(C1)) / C1 + 1 + kLn2f / C1);
    if (w) *w = factor * C2 * (log(cosh(C1 * (1 - q))) / C1 - q + 1 + kLn2f / C1);
    if (g) *g = factor * (-(C2 * tanh(C1 * (1 - q))) - 1) / mh;
  }
}

__device__ __forceinline__ bool kern_wendland(int skf) { return skf >= 4 && skf <= 6; }

__device__ __forceinline__ double kern_grad(int skf, double r, double mh) {
  if (!kern_wendland(skf)) {
    double g;
    kern_other(skf, r, mh, nullptr, &g, nullptr);
    return g;
  }
  if (skf == 4) {
    const double q = r / mh * 1.0, ih = 1.0 / mh;
    const double factor = kC2 * (ih * ih);
    return factor * (20. * (ipow<4>(q) - 3 * ipow<3>(q) + 3 * ipow<2>(q) - q)) / mh;
  }
  if (skf == 5) {
    const double q = r / mh * 1.0, ih = 1.0 / mh;
    const double factor = kC4 * (ih * ih);
    return factor * (-(56 * q * (5 * q + 1) * ipow<5>(1 - q))) / mh;
  }
  return c6_grad(r, mh);
}

__device__ __forceinline__ double kern_w(int skf, double r, double mh) {
  if (!kern_wendland(skf)) {
    double w;
    kern_other(skf, r, mh, &w, nullptr, nullptr);
    return w;
  }
  if (skf == 4) {
    const double q = r / mh * 1.0, ih = 1.0 / mh;
    return kC2 * (ih * ih) * (ipow<4>(1. - q) * (1 + 4 * q));
  }
  if (skf == 5) {
    const double q = r / mh * 1.0, ih = 1.0 / mh;
    return kC4 * (ih * ih) * (ipow<6>(1 - q) * (35 * ipow<2>(q) + 18 * q + 3));
  }
  return c6_w(r, mh);
}

__device__ __forceinline__ double kern_zero(int skf, double mh) {
  if (!kern_wendland(skf)) {
    double kz;
    kern_other(skf, 0.0, mh, nullptr, nullptr, &kz);
    return kz;
  }
  const double ih = 1.0 / mh;
  if (skf == 4) return kC2 * (ih * ih);
  if (skf == 5) return kC4 * (ih * ih) * 3;
  return c6_factor(mh);
}

__device__ __forceinline__ float kern_gradr_f(int skf, float d2, float mh) {
  if (!kern_wendland(skf)) {
    // (the other kernels in FP64: their gradients need not vanish at r = 0)
    const double r = sqrt((double)d2);
    return (float)(kern_grad(skf, r, (double)mh) * kern_scale(skf) / r);
  }
  if (skf == 4 || skf == 5) {
    const float ih = __builtin_amdgcn_rcpf(mh);
    const float q = __builtin_amdgcn_sqrtf(d2) * ih;
    const float ih2 = ih * ih;
    const float omq = 1.0f - q, o2 = omq * omq;
    if (skf == 4) return (-20.0f * kC2f) * (ih2 * ih2) * (o2 * omq);
    return (-56.0f * kC4f) * (ih2 * ih2) * (o2 * o2 * omq) * __builtin_fmaf(5.0f, q, 1.0f);
  }
  return c6_gradr_f(d2, mh);
}

// The boundary kernels' values (boundary_kernel.f90): b_Wendland_C6 reads
// hsml_ij at the boundary-pair index (:491), C2 and C4 scale with b_length
// (:343-346, :417-420).
__device__ __forceinline__ double bkern_w(int skf, double br, double hk) {
  if (!kern_wendland(skf)) {
    // b_Quadratic (boundary_kernel.f90:106-110), b_quintic_spline
    // (:250-272), b_Gaussian (:177-181, also for the Marquis kernels)
    const double sh = kern_scale(skf), q = br / kBLength * sh, t = sh / kBLength;
    if (skf == 1) return 2. / kPi * (t * t) * (3. / 16. * ipow<2>(q) - 3. / 4. * q + 3. / 4.);
    if (skf == 3) {
      const double factor = 15. / (478. * kPi) * (t * t);
      const double a = 3 - q, b = 2 - q, c = 1 - q;
      if (q >= 0. && q <= 1.) return factor * (ipow<5>(a) - 6 * ipow<5>(b) + 15 * ipow<5>(c));
      if (q > 1. && q <= 2.) return factor * (ipow<5>(a) - 6 * ipow<5>(b));
      return factor * ipow<5>(a);
    }
    return 1. / kPi * (t * t) * exp(-ipow<2>(q));
  }
  const double q = br / kBLength * 1.0;
  const double ih = (skf == 4 || skf == 5) ? 1.0 / kBLength : 1.0 / hk;
  if (skf == 4) return kC2 * (ih * ih) * (ipow<4>(1. - q) * (1 + 4 * q));
  if (skf == 5) return kC4 * (ih * ih) * (ipow<6>(1 - q) * (35 * ipow<2>(q) + 18 * q + 3));
  const double omq = 1 - q;
  return kC6 * (ih * ih) * ipow<8>(omq) * (32 * ipow<3>(q) + 25 * ipow<2>(q) + 8 * q + 1);
}

// The boundary kernels' gradient scalar g: bdwdx = -(g (bdx / br)
// scale_h), g = factor inner / b_length with each b_ routine's bracket in its
// evaluation order (boundary_kernel.f90:110-112, 182-183, 257-274, 348-349,
// 422-423, 495-496)
__device__ double bkern_g(int skf, double br, double hk) {
  const double sh = kern_scale(skf), q = br / kBLength * sh, t = sh / kBLength;
  double factor, inner;
  if (skf == 1) {
    factor = 2. / kPi * (t * t);
    inner = 3. / 8. * q - 3. / 4.;
  } else if (skf == 2 || skf == 7 || skf == 8) {
    factor = 1. / kPi * (t * t);
    inner = -(2 * q * exp(-ipow<2>(q)));
  } else if (skf == 3) {
    factor = 15. / (478. * kPi) * (t * t);
    const double a = ipow<4>(3 - q), b = ipow<4>(2 - q), c = ipow<4>(1 - q);
    if (q >= 0. && q <= 1.)
      inner = -(5 * a) + 30 * b - 75 * c;
    else if (q > 1. && q <= 2.)
      inner = -(5 * a) + 30 * b;
    else
      inner = -(5 * a);
  } else if (skf == 4) {
    factor = 3. / (2. * kPi) * (t * t);
    inner = 20. * (ipow<4>(q) - 3 * ipow<3>(q) + 3 * ipow<2>(q) - q);
  } else if (skf == 5) {
    factor = 3. / kPi * (t * t);
    inner = -(56 * q * (5 * q + 1) * ipow<5>(1 - q));
  } else {
    const double ih = 1.0 / hk;
    factor = 78. / (7. * kPi) * (ih * ih);
    inner = -(22. * q * ipow<7>(1 - q) * (16 * ipow<2>(q) + 7. * q + 1));
  }
  return factor * inner / kBLength;
}

__device__ __forceinline__ float bkern_w_f(int skf, float q, float hk) {
  const float omq = 1.0f - q, o2 = omq * omq, o4 = o2 * o2;
  if (skf == 4 || skf == 5) {
    const float ih = (float)(1.0 / kBLength);
    if (skf == 4) return kC2f * (ih * ih) * o4 * __builtin_fmaf(4.0f, q, 1.0f);
    return kC4f * (ih * ih) * (o4 * o2) *
           __builtin_fmaf(__builtin_fmaf(35.0f, q, 18.0f), q, 3.0f);
  }
  const float ih = __builtin_amdgcn_rcpf(hk);
  const float poly = __builtin_fmaf(
      __builtin_fmaf(__builtin_fmaf(32.0f, q, 25.0f), q, 8.0f), q, 1.0f);
  return kC6f * (ih * ih) * (o4 * o4) * poly;
}

// MC_bulk_viscosity and MC_shear_viscosity (bulk_viscosity.f90:113-143,
// shear_viscosity.f90:91-122): principal strain rates, zeta from their sum
// and eta from their difference, each floored at 1e-18 and capped at
// zeta_max P, eta_max P.  The reference declares zeta_max and eta_max but
// never assigns them (data_variables.f90:153-155); its compiled module
// storage holds zero, and so do these caps.  mu = TAN(45 pi / 180.) as the
// reference build folds it, 1 - 2^-53.
constexpr double kMuMC = 0x1.fffffffffffffp-1;
constexpr double kZetaMax = 0.0, kEtaMax = 0.0, kZeroMC = 1.e-18;

__device__ __forceinline__ void mc_viscosity(double e11, double e12, double e22, double P,
                                             double &z, double &et) {
  const double sq = sqrt((e11 - e22) * (e11 - e22) + 4 * (e12 * e12));
  const double e1 = 0.5 * ((e11 + e22) + sq);
  const double e2 = 0.5 * ((e11 + e22) - sq);
  double eI = fabs(e1 + e2);
  if (eI < kZeroMC) eI = kZeroMC;
  double zz = 0.5 * P * (1 + kKT) / eI;
  if (zz > kZetaMax * P) zz = kZetaMax * P;
  double eII = e1 - e2;
  if (fabs(eII) < kZeroMC) eII = kZeroMC;
  double ee = kMuMC * (0.5 * P * (1 + kKT) - zz * (e1 + e2)) / eII;
  if (ee > kEtaMax * P) ee = kEtaMax * P;
  z = zz;
  et = ee;
}

// Compensated (TwoSum) FP32 accumulator.
struct KSum {
  float s = 0.0f, c = 0.0f;
  __device__ __forceinline__ void add(float v) {
    float t = s + v;
    float bp = t - s;
    c += (s - (t - bp)) + (v - bp);
    s = t;
  }
  __device__ __forceinline__ double value() const { return (double)s + (double)c; }
};

// ---------------------------------------------------------------- pass A

// LDS tile of the block's candidate region (FP32 prefilter data; hsml
// pre-multiplied by kHalfPad = sqrt(1.00001) / 2, the prefilter's relative
// margin folded in)
constexpr int kTile = 1280;

// The cell-list runs holding every candidate of a block's particles: the
// block's cell range [cf, cl] shifted by -W, 0 and +W cells, widened by one
// cell on each side; overlapping runs (a block spanning most of a row) are
// merged.  Tile index i < T covers the runs in order.
struct TileRuns {
  int As, Bs, Cs, tB, tC, T;
};
__device__ __forceinline__ TileRuns tile_runs(const int *cs0, int cf, int cl, int W) {
  TileRuns r;
  int Ae, Be, Ce;
  r.As = cs0[cf - W - 1]; Ae = cs0[cl - W + 2];
  r.Bs = cs0[cf - 1];     Be = cs0[cl + 2];
  r.Cs = cs0[cf + W - 1]; Ce = cs0[cl + W + 2];
  if (Ae >= r.Bs) { r.Bs = r.As; Be = max(Ae, Be); r.As = Ae = 0x7fffffff; }
  if (Be >= r.Cs) { r.Cs = min(r.Bs, r.Cs); Ce = max(Be, Ce); r.Bs = Be = 0x7fffffff; }
  const int lA = Ae - r.As, lB = Be - r.Bs, lC = Ce - r.Cs;
  r.tB = lA;
  r.tC = lA + lB;
  r.T = lA + lB + lC;
  return r;
}
__device__ __forceinline__ int tile_slot(const TileRuns &r, int i) {
  return i < r.tB ? r.As + i : (i < r.tC ? r.Bs + (i - r.tB) : r.Cs + (i - r.tC));
}
__device__ __forceinline__ int tile_index(const TileRuns &r, int sp) {
  return sp >= r.Cs ? sp - r.Cs + r.tC : (sp >= r.Bs ? sp - r.Bs + r.tB : sp - r.As);
}
constexpr float kPad = 1.000005f;
// candidate-list skin, relative to MINVAL(hsml)
constexpr float kHalfPad = 0.5f * kPad;

// XSW: instances with switch values beyond the reference's defaults may be
// present (summation density, stress_2, the Hibler or Gutfraind pressure:
// their code is compiled in; each instance still follows its own values)
// BA: candidates whose loads go out together in the exact test (a tuning
// parameter: the results do not depend on it)
template <bool MIXED, bool XSW, int BA>
__global__ void __launch_bounds__(kTPB) __attribute__((amdgpu_waves_per_eu(4)))
k_passA(DevState d, int buf, int want_diag) {
  constexpr int kBatch = BA;
  __shared__ float4 tile[kTile];
  __shared__ unsigned short lst[kNbrMax * kTPB];
  if (blockIdx.x >= (unsigned)d.nblk) return;
  // The block entry carries the instance's current state set: the
  // particle's own loads go out right after it, together with the loads of
  // the instance records, instead of after them.
  const BlockEntry be = d.blk[blockIdx.x];
  const int k = be.inst;
  // the evaluated state, by slot (= list position)
  const int ss = be.cs, sb = 2 * ss + buf;
  const double *X = d.x[sb], *Y = d.y[sb], *HS = d.hsml[sb], *RS = d.rho[sb];
  const double *VX = d.vx[sb], *VY = d.vy[sb];
  const double *MS = (XSW && buf && be.tmass) ? d.masst[ss] : d.mass[ss];
  const int pos0 = be.pos;                // first slot of the block
  const int nloc = be.nloc;
  const bool valid = (int)threadIdx.x < nloc;
  const int self = pos0 + min((int)threadIdx.x, nloc - 1);
  const size_t N = (size_t)d.ntot;
  const int c = d.cellid[ss][self];
  const double xi = X[self], yi = Y[self], hi = HS[self];
  const double vxi = VX[self], vyi = VY[self];
  // kept candidate lists: the count and the first kBatch entries (entries
  // past the count are never dereferenced; unused when the lists are
  // rebuilt in this evaluation)
  int vcount = d.vnc[self], nx[kBatch];
#pragma unroll
  for (int u = 0; u < kBatch; ++u) nx[u] = d.vnb[u * N + self];
  // the rheology's thickness and concentration, loaded with the rest (the
  // FP64 variant, short of registers, loads them where they are used)
  double hth_i = 0., A_i = 0., rho_own = 0., m_own = 0.;
  if (MIXED) {
    hth_i = d.h[sb][self];
    A_i = d.A[sb][self];
    rho_own = RS[self];
    m_own = MS[self];
  }
  // the previous evaluation's pair counts, loaded here with the rest: the
  // compiler cannot move a load above the list stores below
  const int oc_old = d.own_cnt[self];
  const InstDyn &y = d.dyn[k];
  // particles whose owned pairs can rank below the boundary kernel's limit
  // (the head of the pair list) also compare the owned neighbours themselves
  const bool head = self - be.pos + be.first <= y.rstop;
  const unsigned osig_old = head ? (unsigned)d.osig[self] : 0u;
  if (blockIdx.x == 0 && threadIdx.x == 0) *d.rebuild = 0;   // the rebuild kernels have run
  if (inst_off(y)) return;                // whole block: blocks never straddle instances
  const InstMeta &m = d.meta[k];
  const int W = y.W;
  const int *cs0 = d.cell_start + m.c_off;
  // the cell's boundary stamp (with boundaries): issued before the gathers
  const int bfl = m.nb > 0 ? d.bflag[m.c_off + c] : -1;
  // Candidate lists: built with the prefilter below (every candidate within
  // mhsml + skin), they still cover the pair set while the cell lists are
  // unchanged (same cells, same visiting order) and every pair's distance
  // and mhsml moved by less than the skin: 2 max|r - r_built| +
  // max|hsml - hsml_built| < skin, bounded by finalize from the largest
  // speed and hsml change of every evaluation since.  Then the prefilter
  // is skipped.
  const bool vok = y.rb == 0 && !lists_stale(y);
  // search statistics of the instance: evaluations, and those that rebuilt
  // the cell lists / the candidate lists
  if (be.first == 0 && threadIdx.x == 0) {
    InstDyn &yw = d.dyn[k];
    yw.n_eval += 1;
    if (y.rb) yw.n_cell += 1;
    if (!vok) yw.n_cand += 1;
  }
  const float skin = d.skin * (float)y.hmin;
  // The candidates of every particle of the block lie in three runs of the
  // cell lists: its cell range shifted by -W, 0 and +W cells, widened by one
  // cell on each side.  Overlapping runs (a block spanning most of a row)
  // are merged.  The runs' prefilter data goes to LDS in one round of
  // coalesced loads by the whole block; a block whose region does not fit
  // takes the exact rescan path instead.
  TileRuns tr{};
  bool tiled = false;
  if (!vok) {
    tr = tile_runs(cs0, d.cellid[ss][pos0], d.cellid[ss][pos0 + nloc - 1], W);
    tiled = tr.T <= kTile;      // list entries are tile indices (< kTile)
    if (tiled) {
      for (int i = threadIdx.x; i < tr.T; i += kTPB) {
        const int s = tile_slot(tr, i);
        float4 v = d.pf[s];
        v.z *= kHalfPad;
        tile[i] = v;
      }
    }
    __syncthreads();
    // lists built in this evaluation carry this skin (any block may say so)
    if (threadIdx.x == 0) d.dyn[k].vskin_next = skin;
  }
  if (!valid) return;
  // waves with large supports (more neighbours) get issue priority: they
  // set the kernel's length
  if (__any(hi > 1.4 * y.hmin)) __builtin_amdgcn_s_setprio(2);
  const int cy = c / W, cx = c - cy * W;
  const double rhoi = XSW ? RS[self] : 0.0;
  const Ranges R = neighbour_ranges(d.cell_start + m.c_off, W, cx, cy);

  const int p = self;
  // summation density (density.f90:41-100) on this evaluation's kernel
  // values, accumulated in the same per-particle order as the reference.
  // FP64 in both precision modes: the density sets the next smoothing
  // length, and through it the pair predicate d <= mhsml, which is not
  // continuous -- an FP32 error there flips borderline pairs.
  const bool sumd = XSW && m.den_alg == 1;
  // the Hernquist-Katz viscosity's velocity divergence (regularization.f90
  // :148-167, each term over rho_i; FP64 in both precision modes)
  const bool hk = XSW && m.avis == 1;
  double div = 0., g1 = 0., g2 = 0., g3 = 0., g4 = 0.;
  double num = 0., den = 0., dvhk = 0.;
  KSum fdiv, f1, f2, f3, f4;
  // the pair predicate; d itself only where the FP64 sums use it
  auto support = [&](double d2, double mh, double &r) {
    return (MIXED && !sumd && !hk) ? in_support_nr(d2, mh) : in_support(d2, mh, r);
  };
  // kernel gradient, vector_divergence and vector_gradient of one pair
  const int skf = XSW ? m.skf : 6;
  const double ksc = XSW ? kern_scale(skf) : 1.0;
  auto pair = [&](double dx, double dy, double r, double mh, double mq, double rq, double vxq,
                  double vyq) {
    const double dv0 = vxi - vxq, dv1 = vyi - vyq;
    if (sumd) {
      const double w = kern_w(skf, r, mh);
      den = den + mq * w / rq;
      num = num + mq * w;
    }
    if (hk) {
      const double gg = kern_grad(skf, r, mh);
      const double dw0 = gg * (dx / r) * ksc, dw1 = gg * (dy / r) * ksc;
      dvhk = dvhk - mq * (dv0 * dw0 + dv1 * dw1) / rhoi;
    }
    if (!MIXED) {
      const double gg = kern_grad(skf, r, mh);
      const double dw0 = gg * (dx / r) * ksc, dw1 = gg * (dy / r) * ksc;
      const double w1 = dv0 * dw0, w2 = dv1 * dw0, w3 = dv0 * dw1, w4 = dv1 * dw1;
      div = div - mq * (w1 + w4);
      g1 = g1 - mq * w1 / rq;
      g2 = g2 - mq * w2 / rq;
      g3 = g3 - mq * w3 / rq;
      g4 = g4 - mq * w4 / rq;
    } else {
      const float fx = (float)dx, fy = (float)dy;
      const float gr = kern_gradr_f(skf, fx * fx + fy * fy, (float)mh);
      const float dw0 = fx * gr, dw1 = fy * gr;
      const float a0 = (float)dv0, a1 = (float)dv1;
      const float w1 = a0 * dw0, w2 = a1 * dw0, w3 = a0 * dw1, w4 = a1 * dw1;
      const float mf = (float)mq, mr = mf * __builtin_amdgcn_rcpf((float)rq);
      fdiv.add(-(mf * (w1 + w4)));
      f1.add(-(mr * w1));
      f2.add(-(mr * w2));
      f3.add(-(mr * w3));
      f4.add(-(mr * w4));
    }
  };
  // 1) FP32 prefilter of the 3x3 block into a candidate list (from the
  //    LDS tile); 2) exact FP64 test and pair arithmetic over the
  //    candidates, which also writes the neighbour list.  Splitting the two
  //    keeps the expensive arithmetic off the ~90 % of candidates that fail.
  int nI = 0, nX = 0, nn = 0, nc = 0;
  unsigned osg = 0u;          // hash of the owned neighbours' slots, in order
  if (vok) {
    nc = vcount;
  } else if (!tiled) {
    nc = kNbrMax + 1;
  } else {
    // tile index of a list position (every position of the block's 3x3
    // blocks lies in one of the runs)
    auto tix = [&](int sp) { return tile_index(tr, sp); };
    const int tself = tix(self);
    const float4 pi4 = tile[tself];
    const float hif = (float)hi;
    const float slack = y.pf_slack;
    // Cell culling: a neighbour cell is scanned only if its nearest point
    // is within the largest support a pair with one of its members can
    // have, (hsml_i + max hsml of the cell) / 2 (cells span
    // [(c - 1/2) bin, (c + 1/2) bin) in the prefilter's frame).  The
    // scanned ranges shrink at their ends only, so the visiting order and
    // the list are those of the full block.
    const float binf = (float)y.bin;
    const float dxl = fmaxf(pi4.x - ((float)cx - 0.5f) * binf - slack, 0.0f);
    const float dxr = fmaxf(((float)cx + 0.5f) * binf - pi4.x - slack, 0.0f);
    const float dyb = fmaxf(pi4.y - ((float)cy - 0.5f) * binf - slack, 0.0f);
    const float dyt = fmaxf(((float)cy + 0.5f) * binf - pi4.y - slack, 0.0f);
    const unsigned long long *ch = d.cellh + m.c_off;
    const unsigned stampu = (unsigned)y.stamp;
    const int *cs = d.cell_start + m.c_off;
    const int rm = (cy - 1) * W + cx, r0 = cy * W + cx, rp = (cy + 1) * W + cx;
    auto need = [&](int c, float ex, float ey) {
      const unsigned long long hc = ch[c];
      const float hmax = (unsigned)(hc >> 32) == stampu ? __uint_as_float((unsigned)hc) : 0.0f;
      const float bnd = __builtin_fmaf(hif + hmax, 0.5f, slack + skin);
      return __builtin_fmaf(ex, ex, ey * ey) <= bnd * bnd * 1.00001f;
    };
    // The four segments, in visiting order: row below (cells rm-1, rm,
    // rm+1 -> the span from the first to the last needed), left / own /
    // right, up / up-right, up-left.
    int a0 = 0, b0 = 0, a2 = 0, b2 = 0, a3 = 0, b3 = 0;
    {
      const bool l = need(rm - 1, dxl, dyb), c0 = need(rm, 0.0f, dyb), r = need(rm + 1, dxr, dyb);
      const int f = l ? 0 : c0 ? 1 : r ? 2 : 3, e = r ? 3 : c0 ? 2 : l ? 1 : 0;
      if (f < e) { a0 = cs[rm - 1 + f]; b0 = cs[rm - 1 + e]; }
    }
    const int a1 = need(r0 - 1, dxl, 0.0f) ? R.a[1] : R.a[2];
    const int b1 = need(r0 + 1, dxr, 0.0f) ? R.b[3] : R.b[2];
    {
      const bool u = need(rp, 0.0f, dyt), ur = need(rp + 1, dxr, dyt);
      const int f = u ? 0 : ur ? 1 : 2, e = ur ? 2 : u ? 1 : 0;
      if (f < e) { a2 = cs[rp + f]; b2 = cs[rp + e]; }
      if (need(rp - 1, dxl, dyt)) { a3 = R.a[5]; b3 = R.a[4]; }
    }
    // One loop over the concatenated segments: a wavefront then runs as
    // long as its longest lane's total, not the sum of per-segment maxima.
    // j -> tile index j + o_k on segment k.
    const int e0 = b0 - a0, e1 = e0 + (b1 - a1), e2 = e1 + (b2 - a2), e3 = e2 + (b3 - a3);
    const int o0 = e0 > 0 ? tix(a0) : 0;
    const int o1 = tix(a1) - e0;
    const int o2 = (b2 > a2 ? tix(a2) : 0) - e1;
    const int o3 = (b3 > a3 ? tix(a3) : 0) - e2;
    // (mhsml + slack + skin) sqrt(1.00001) = hzi + the candidate's scaled hsml
    const float hzi = __builtin_fmaf(slack + skin, kPad, pi4.z);
    for (int j = 0; j < e3; ++j) {
      const int ti = j + (j < e0 ? o0 : (j < e1 ? o1 : (j < e2 ? o2 : o3)));
      const float4 q = tile[ti];
      // rejects only candidates certainly beyond mhsml
      const float fx = pi4.x - q.x, fy = pi4.y - q.y;
      const float fm = hzi + q.z;
      const bool hit =
          (ti != tself) & (__builtin_fmaf(-fm, fm, __builtin_fmaf(fx, fx, fy * fy)) <= 0.0f);
      if (hit) {
        if (nc < kNbrMax) lst[nc * kTPB + threadIdx.x] = (unsigned short)ti;
        ++nc;
      }
    }
  }
  // list entry (tile index) -> list position
  auto lpos = [&](int ti) { return tile_slot(tr, ti); };
  // list position of candidate j
  auto cand = [&](int j) {
    return vok ? d.vnb[j * N + self] : lpos((int)lst[j * kTPB + threadIdx.x]);
  };
  if (!vok) {
    // this evaluation's candidate lists, kept for the next ones
    d.vnc[self] = nc;
    if (nc <= kNbrMax)
      for (int j = 0; j < nc; ++j) d.vnb[j * N + self] = cand(j);
  }
  // kernel_zero(i) (kernel.f90:503): the factor of the last pair in the
  // reference's order that has this particle first, i.e. its last owned
  // pair; particles that own none keep the value of an earlier evaluation.
  // Owned candidates form the tail of the visiting order, so a backward
  // walk finds it.  It must be known before the density sums start.
  double kzi = 0.;
  if (sumd) {
    kzi = d.kz[m.p_off + d.orig[ss][self]];
    bool found = false;
    if (nc <= kNbrMax) {
      for (int j = nc - 1; j >= 0 && !found; --j) {
        const int t = cand(j);
        if (pair_kind(R, t, self) == 0) break;
        const double dx = xi - X[t], dy = yi - Y[t];
        const double mh = (hi + HS[t]) / 2.;
        double r;
        if (in_support(dx * dx + dy * dy, mh, r)) {
          kzi = kern_zero(skf, mh);
          found = true;
        }
      }
    } else {
      for (int g = 5; g >= 2 && !found; --g) {
        for (int t = R.b[g] - 1; t >= R.a[g] && !found; --t) {
          if (g == 2 && t <= self) break;
          const double dx = xi - X[t], dy = yi - Y[t];
          const double mh = (hi + HS[t]) / 2.;
          double r;
          if (in_support(dx * dx + dy * dy, mh, r)) {
            kzi = kern_zero(skf, mh);
            found = true;
          }
        }
      }
    }
    const double mi = MS[self];
    num = kzi * mi;
    den = kzi * mi / rhoi;
  }
  if (nc <= kNbrMax) {
    for (int j0 = 0; j0 < nc; j0 += kBatch) {
      int sj[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u)
        sj[u] = j0 + u < nc ? (vok ? nx[u] : cand(j0 + u)) : self;
      // positions and pair data of the batch in one round of loads
      double xq[kBatch], yq[kBatch], hq[kBatch], mq[kBatch], rq[kBatch], uq[kBatch], wq[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        xq[u] = X[sj[u]];
        yq[u] = Y[sj[u]];
        hq[u] = HS[sj[u]];
        mq[u] = MS[sj[u]];
        rq[u] = RS[sj[u]];
        uq[u] = VX[sj[u]];
        wq[u] = VY[sj[u]];
      }
      // kept lists: the next batch's entries go out with this batch's data
      // (past the end, the last entry again)
      if (vok) {
#pragma unroll
        for (int u = 0; u < kBatch; ++u) nx[u] = d.vnb[min(j0 + kBatch + u, nc - 1) * N + self];
      }
      bool hit[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const double dx = xi - xq[u], dy = yi - yq[u];
        const double mh = (hi + hq[u]) / 2.;
        double r = 0.;
        hit[u] = j0 + u < nc && support(dx * dx + dy * dy, mh, r);
        if (!hit[u]) continue;
        const int kind = pair_kind(R, sj[u], self);
        nI += kind == 1;
        nX += kind == 2;
        if (kind != 0) osg = osg * 0x9e3779b1u + (unsigned)sj[u];
        pair(dx, dy, r, mh, mq[u], rq[u], uq[u], wq[u]);
      }
      // The batch's list entries: one store per candidate, hit or not, at
      // the next free position -- a miss's entry is overwritten by the next
      // hit or lies past the list's end (row kNbrMax is spare).  With a
      // fixed number of stores per batch the next batch's wait for its
      // list entries, issued before these, does not wait for the stores to
      // complete (a data-dependent store count made it wait for all).
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        d.nbr[nn * N + self] = sj[u];
        nn += hit[u] ? 1 : 0;
      }
    }
  } else {
    nn = 0;
    int ntrue = 0;
#pragma unroll
    for (int g = 0; g < 6; ++g) {
      for (int s = R.a[g]; s < R.b[g]; ++s) {
        if (s == self) continue;
        const double dx = xi - X[s], dy = yi - Y[s];
        const double mh = (hi + HS[s]) / 2.;
        double r = 0.;
        if (!support(dx * dx + dy * dy, mh, r)) continue;
        const int kind = pair_kind(R, s, self);
        nI += kind == 1;
        nX += kind == 2;
        if (kind != 0) osg = osg * 0x9e3779b1u + (unsigned)s;
        ++ntrue;
        pair(dx, dy, r, mh, MS[s], RS[s], VX[s], VY[s]);
      }
    }
    nn = ntrue;
  }
  // list length, or kNbrMax + 1 when it was not stored (the other gathers
  // then rescan the 3x3 block)
  d.nnb[self] = nc <= kNbrMax ? nn : kNbrMax + 1;
  if (MIXED) {
    div = fdiv.value();
    g1 = f1.value(); g2 = f2.value(); g3 = f3.value(); g4 = f4.value();
  }
  if (sumd) {
    d.kz[m.p_off + d.orig[ss][self]] = kzi;
    d.dnum[self] = num;
    d.dsum[self] = num / den;
  }
  if (hk) {
    // no viscosity for diverging flow; lv (regularization.f90:170-176) with
    // the sound speed the switches leave (c_yang of the density, or zero)
    const double dv = dvhk >= 0. ? 0. : dvhk;
    const double ci = m.needc ? sound_speed(rhoi) : 0.0;
    const double t = fabs(dv) * hi;
    d.lv[self] = (kAlphaPi * hi * ci * fabs(dv) + kBetaPi * (t * t)) * rhoi;
  }
  // boundary pairs per (boundary particle, cell of its 3x3 block): the
  // counts bucket_search's boundary loop finds there (nnps.f90:453-492),
  // gathered from the particle side; k_ranks turns them into ranks
  if (m.nb > 0 && bfl == y.lstamp) {
    const int2 run = d.brun[m.c_off + c];
    const int *bl = d.blist + 9 * (size_t)m.b_off;
    // entries in groups of four: their loads, then the boundary particles'
    // positions, go out together
    for (int e0 = run.x; e0 < run.y; e0 += 4) {
      int ent[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) ent[u] = e0 + u < run.y ? bl[e0 + u] : -1;
      double bx[4], by[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int bg = m.b_off + (ent[u] < 0 ? 0 : (ent[u] & 0xffffff));
        bx[u] = d.xb[bg];
        by[u] = d.yb[bg];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (ent[u] < 0) continue;
        const int bg = m.b_off + (ent[u] & 0xffffff);
        const double bdx = bx[u] - xi, bdy = by[u] - yi;
        if (sqrt(bdx * bdx + bdy * bdy) <= kBLength)
          atomicAdd(&d.bcnt[(size_t)bg * 9 + (ent[u] >> 24)], 1);
      }
    }
  }
  if (nI > 1023 || nX > 1023) atomicOr(&d.dyn[k].err, kErrCapacity);
  const int oc_new = (nI & 1023) | ((nX & 1023) << 10) | (min(nn, 4095) << 20);
  // the pairs k_ranks keeps for the ranks below its limit are the owned
  // pairs of the list positions up to its stop: a change of their counts or
  // of the owned neighbours has them found again
  if (head) d.osig[self] = (int)osg;
  if (__any(head && (((oc_new ^ oc_old) & 0xfffff) != 0 || osg != osig_old)) &&
      ((int)threadIdx.x & 63) == 0)
    atomicOr(&d.dyn[k].rchg, 1);
  d.own_cnt[self] = oc_new;

  // strain rate (single_step.f90:80-84); with viscosity and the pressure
  // force off it is never set and keeps initialize's zero
  const bool strain = !(XSW && m.nostrain);
  const double e11 = strain ? g1 : 0.0, e22 = strain ? g4 : 0.0;
  const double e12 = strain ? 0.5 * (g2 + g3) : 0.0;
  // Hibler P (internal_stress.f90:24), elliptical zeta / eta
  // (bulk_viscosity.f90:65-76, shear_viscosity.f90:57), Kreyscher
  // replacement pressure (eos.f90:141-147), stress tensor
  // (stress_tensor.f90:31-39).  FP64 in both precision modes: the
  // delta -> delta_min switch is decided here.
  const double hth = MIXED ? hth_i : d.h[sb][self], Ac = MIXED ? A_i : d.A[sb][self];
  const int eos = XSW ? m.eos_alg : 3;
  const double rho_i = MIXED ? rho_own : (eos == 2 ? RS[self] : 0.0), m_i = m_own;
  const double em2 = 1.0 / (kE * kE);
  const double P = kPStar * hth * exp(-(kCEos * (1 - Ac)));
  const double delta = sqrt((e11 * e11 + e22 * e22) * (1 + em2) + 4 * em2 * (e12 * e12) +
                            2 * e11 * e22 * (1 - em2));
  const double dcl = fabs(delta) < kDeltaMin ? kDeltaMin : delta;
  double z = 0.5 * P * (1 + kKT) / dcl;
  double et = z / (kE * kE);
  if (XSW && m.rheology == 2) mc_viscosity(e11, e12, e22, P, z, et);
  if (XSW && m.novisc) {
    // the viscosity block (internal_stress.f90:20-31) is skipped: zeta and
    // eta keep initialize's zero
    z = 0.0;
    et = 0.0;
  }
  double pr;
  if (eos == 1) {
    pr = P;                                    // hibler_eos (eos.f90:59)
  } else if (eos == 0) {
    pr = 0.0;                                  // neither eos nor the viscosity block
  } else if (eos == 2) {
    // gutfraind_eos (eos.f90:94-102), rho_0 = rho_ice, negative
    // pressures set to zero
    const double gam = 1.9e5 * hth * exp(-(20.0 * (1 - Ac)));
    pr = gam * (rho_i - kRhoIce) / (kRhoMaxEos - rho_i);
    if (pr < 0.) pr = 0.;
  } else {
    pr = z * 2 * delta / (1 + kKT);            // kreyscher_eos (eos.f90:147)
  }
  const double s11 = (z + et) * e11 + (z - et) * e22 - pr * (1 - kKT) / 2;
  const double s12 = 2 * et * e12;
  const double s22 = (z + et) * e22 + (z - et) * e11 - pr * (1 - kKT) / 2;
  if (!MIXED) {
    ((double *)d.sig11)[self] = s11;
    ((double *)d.sig12)[self] = s12;
    ((double *)d.sig22)[self] = s22;
  } else {
    // pass B's FP32 stress sums read one 16-byte record per neighbour:
    // stress_1 {sigma / rho^2, m}, stress_2 {sigma, m / rho}
    if (!XSW || m.pa_sph != 2) {
      const double ir2 = 1.0 / (rho_i * rho_i);
      d.srec[self] = make_float4((float)(s11 * ir2), (float)(s12 * ir2), (float)(s22 * ir2),
                                 (float)m_i);
    } else {
      d.srec[self] = make_float4((float)s11, (float)s12, (float)s22, (float)(m_i / rho_i));
    }
  }
  ((double *)d.div)[self] = div;
  if (want_diag) {
    const int N = d.ntot;
    d.diag[p] = e11;
    d.diag[N + p] = e12;
    d.diag[2 * N + p] = e22;
    d.diag[3 * N + p] = z;
    d.diag[4 * N + p] = et;
    d.diag[5 * N + p] = pr;
    if (!MIXED) {
      d.sig_out[p] = s11;
      d.sig_out[N + p] = s12;
      d.sig_out[2 * N + p] = s22;
    } else {
      d.sig_out[p] = (double)(float)s11;
      d.sig_out[N + p] = (double)(float)s12;
      d.sig_out[2 * N + p] = (double)(float)s22;
    }
  }
}

// ------------------------------------------------------- pair-rank store

// Per cell: within-cell offsets of each member's owned intra and cross
// pairs, and the cell's pair total.
__global__ void __launch_bounds__(kTPB) k_pairtot(DevState d) {
  int c = blockIdx.x * kTPB + threadIdx.x;
  if (c >= d.ctot) return;
  int s0 = d.cell_start[c], s1 = d.cell_start[c + 1];
  int I = 0, X = 0;
  for (int s = s0; s < s1; ++s) {
    int oc = d.own_cnt[s];
    d.moff[s] = I | (X << 16);
    I += oc & 1023;
    X += (oc >> 10) & 1023;
  }
  d.pair_count[c] = I + X;
  d.cell_intra[c] = I;
}

// hsml_ij of the particle pairs ranked below min(bnip, store size): the
// entries b_Wendland_C6 reads at the boundary-pair index
// (boundary_kernel.f90:491); entries at or above nip keep what an earlier
// evaluation wrote, as the reference's module array does.  The ranks are
// the pair prefix over the cell lists (per cell: all intra pairs, then all
// cross pairs, each in member order), and only the pairs ranked below
// min(bnip, store size) are read by the boundary kernel -- a short head of
// the list.  One block per instance finds them, streaming the prefix over
// list positions until it passes that limit, and keeps each such pair's
// owner and partner slots.  While no rebuild moves the slots and the head
// particles keep their owned neighbours (pass A compares them), the pairs
// at those ranks stay the same, and each step only refreshes their hsml_ij
// from the new smoothing lengths.
constexpr int kRankThreads = 1024;

__device__ void rank_store(const DevState &d, const InstDyn &y, const InstMeta &m, int ss,
                           int buf, int t, int lim, int gendI, int gendX);

__global__ void __launch_bounds__(kRankThreads) k_ranks(DevState d, int buf) {
  __shared__ int sh_stop;
  if (blockIdx.x >= (unsigned)d.ninst) return;
  const int k = blockIdx.x;
  const InstDyn &y = d.dyn[k];
  const InstMeta &m = d.meta[k];
  if (m.nb == 0 || inst_off(y)) return;
  // 0) boundary-pair ranks: per boundary particle the prefix over its 9
  //    cells of the counts pass A gathered, then the prefix over boundary
  //    particles (= rank of its first pair) and bnip; the counters are
  //    cleared for the next evaluation
  int bnip = 0;
  {
    const int per = (m.nb + kRankThreads - 1) / kRankThreads;
    const int b0 = threadIdx.x * per;
    int tsum = 0;
    for (int b = b0; b < b0 + per && b < m.nb; ++b) {
      const size_t bg = (size_t)(m.b_off + b);
      int c9[9];
#pragma unroll
      for (int s9 = 0; s9 < 9; ++s9) c9[s9] = d.bcnt[bg * 9 + s9];
      int run = 0;
#pragma unroll
      for (int s9 = 0; s9 < 9; ++s9) {
        d.bcnt[bg * 9 + s9] = 0;
        d.bpre[bg * 9 + s9] = run;
        run += c9[s9];
      }
      d.bcount[bg] = run;
      tsum += run;
    }
    int ex = block_exclusive_scan<kRankThreads>(tsum, &bnip);
    for (int b = b0; b < b0 + per && b < m.nb; ++b) {
      d.bbase[m.b_off + b] = ex;
      ex += d.bcount[m.b_off + b];
    }
    if (threadIdx.x == 0) {
      d.dyn[k].bnip = bnip;
      if (bnip > m.q_cap) atomicOr(&d.dyn[k].err, kErrCapacity);
    }
  }
  const int lim = min(bnip, m.q_cap);
  if (lim <= 0) {
    if (threadIdx.x == 0) {
      d.dyn[k].rlim = 0;
      d.dyn[k].rfresh = 0;
    }
    return;
  }
  const int ss = cur_set(y);
  int2 *HP = d.hpair + m.q_off;
  // 1) the pairs below the limit are those of the last search while the
  //    slots, the limit and the head's owned neighbours are unchanged (pass
  //    A flags a change); their hsml_ij changes only with hsml, at the end of
  //    a step -- the corrector's half-step state has the predictor's hsml.
  //    Every rank below the limit is the index of exactly one boundary pair,
  //    so pass B, which reads each of them once, refreshes the entry from
  //    the pair's slots (rfresh) in the evaluations of a step's start state.
  if (y.rb == 0 && y.rchg == 0 && lim == y.rlim) {
    if (threadIdx.x == 0) d.dyn[k].rfresh = buf == 0;
    return;
  }
  // 2) otherwise search again: no pair at any rank yet
  for (int r = threadIdx.x; r < lim; r += kRankThreads) HP[r] = make_int2(-1, -1);
  const int *cs = d.cell_start + m.c_off;
  const int p0 = cs[0], np = m.n;
  int2 *G = d.gpre + p0;       // per position: exclusive prefixes of intra, cross counts
  if (threadIdx.x == 0) sh_stop = np;
  __syncthreads();
  //    prefixes up to the end of the last cell whose first pair ranks below lim
  constexpr int kPer = 4;      // consecutive positions per thread
  int baseI = 0, baseX = 0;
  for (int c0 = 0; c0 < np && c0 <= sh_stop; c0 += kRankThreads * kPer) {
    const int t0 = c0 + (int)threadIdx.x * kPer;
    int vI[kPer], vX[kPer], sI = 0, sX = 0;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int oc = t0 + u < np ? d.own_cnt[p0 + t0 + u] : 0;
      vI[u] = oc & 1023;
      vX[u] = (oc >> 10) & 1023;
      sI += vI[u];
      sX += vX[u];
    }
    int totI = 0, totX = 0;
    int gI = baseI + block_exclusive_scan<kRankThreads>(sI, &totI);
    int gX = baseX + block_exclusive_scan<kRankThreads>(sX, &totX);
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int t = t0 + u;
      if (t < np) {
        G[t] = make_int2(gI, gX);
        // the last position whose prefix is below lim: its cell is the
        // last one needed
        if (gI + gX < lim && gI + gX + vI[u] + vX[u] >= lim)
          sh_stop = cs[d.cellid[ss][p0 + t] + 1] - p0;
      }
      gI += vI[u];
      gX += vX[u];
    }
    baseI += totI;
    baseX += totX;
    __syncthreads();
  }
  //    then each position's owned pairs below lim: their slots and hsml_ij
  const int stop = sh_stop;
  for (int t = threadIdx.x; t < stop; t += kRankThreads)
    rank_store(d, y, m, ss, buf, t, lim, baseI, baseX);
  if (threadIdx.x == 0) {
    InstDyn &yw = d.dyn[k];
    yw.rfresh = 0;      // the search wrote the entries
    yw.rchg = 0;
    yw.rlim = lim;
    yw.rstop = stop;
    yw.gendI = baseI;   // prefixes at np, when the loop reached it
    yw.gendX = baseX;
  }
}

// The hsml_ij of the pairs list position t owns, where they rank below lim
// (G: pair-rank prefixes, valid up to the end of the last needed cell).
__device__ void rank_store(const DevState &d, const InstDyn &y, const InstMeta &m, int ss,
                           int buf, int t, int lim, int gendI, int gendX) {
  const int *cs = d.cell_start + m.c_off;
  const int p0 = m.p_off, np = m.n, s = p0 + t;
  const int2 *G = d.gpre + p0;
  const size_t N = (size_t)d.ntot;
  const int W = y.W;
  const int c = d.cellid[ss][s];
  const int f = cs[c] - p0, e = cs[c + 1] - p0;
  const int2 Gf = G[f], Gt = G[t];
  const int2 Ge = e < np ? G[e] : make_int2(gendI, gendX);
  const int base = Gf.x + Gf.y;
  if (base >= lim) return;
  const int oc = d.own_cnt[s];
  const int nI = oc & 1023, nX = (oc >> 10) & 1023;
  int rI = base + (Gt.x - Gf.x);
  int rX = base + (Ge.x - Gf.x) + (Gt.y - Gf.y);
  if (!((nI > 0 && rI < lim) || (nX > 0 && rX < lim))) return;
  const double *HS = d.hsml[2 * ss + buf];
  const double hi = HS[s];
  const int nn = d.nnb[s];
  if (nn <= kNbrMax) {
    // the owned pairs are the tail of the list: nI intra, then nX cross
    for (int j = nn - nI - nX; j < nn; ++j) {
      const int q = d.nbr[j * N + s];
      int &rank = (j < nn - nX) ? rI : rX;
      if (rank < lim) {
        d.hq[m.q_off + rank] = (hi + HS[q]) / 2.;
        d.hpair[m.q_off + rank] = make_int2(s, q);
      }
      ++rank;
    }
    return;
  }
  const double *X = d.x[2 * ss + buf], *Y = d.y[2 * ss + buf];
  const int cy = c / W, cx = c - cy * W;
  const double xi = X[s], yi = Y[s];
  const Ranges R = neighbour_ranges(cs, W, cx, cy);
  for (int g = 2; g < 6; ++g) {
    for (int q = R.a[g]; q < R.b[g]; ++q) {
      if (g == 2 && q <= s) continue;
      const double dx = xi - X[q], dy = yi - Y[q];
      const double mh = (hi + HS[q]) / 2.;
      double r;
      if (!in_support(dx * dx + dy * dy, mh, r)) continue;
      int &rank = (g == 2) ? rI : rX;
      if (rank < lim) {
        d.hq[m.q_off + rank] = mh;
        d.hpair[m.q_off + rank] = make_int2(s, q);
      }
      ++rank;
    }
  }
}

// ---------------------------------------------------------------- pass B

__device__ __forceinline__ void wind_at(const DevState &d, const InstMeta &m, double px,
                                        double py, double &u, double &v) {
  if (m.wind_mode != 1) {
    u = m.wind_u;
    v = m.wind_v;
    return;
  }
  const int nx = m.g_nx, ny = m.g_ny;
  const double fx = (px - m.g_x0) / m.g_dx, fy = (py - m.g_y0) / m.g_dy;
  const int i0 = (int)fmin(fmax(floor(fx), 0.0), (double)(nx - 2));
  const int j0 = (int)fmin(fmax(floor(fy), 0.0), (double)(ny - 2));
  const double tx = fmin(fmax(fx - (double)i0, 0.0), 1.0);
  const double ty = fmin(fmax(fy - (double)j0, 0.0), 1.0);
  const double *U = d.grid + m.g_off;
  const double *V = U + (long)nx * ny;
  const long r0 = (long)j0 * nx, r1 = r0 + nx;
  const double u0 = (1.0 - tx) * U[r0 + i0] + tx * U[r0 + i0 + 1];
  const double u1 = (1.0 - tx) * U[r1 + i0] + tx * U[r1 + i0 + 1];
  const double v0 = (1.0 - tx) * V[r0 + i0] + tx * V[r0 + i0 + 1];
  const double v1 = (1.0 - tx) * V[r1 + i0] + tx * V[r1 + i0 + 1];
  u = (1.0 - ty) * u0 + ty * u1;
  v = (1.0 - ty) * v0 + ty * v1;
}

// phase: 0 = predictor (state S -> T), 1 = corrector (RHS on T, update S in
// place), 2 = single evaluation (tendencies to d.tend), 3 = Euler step (the
// tendencies to d.tend, which k_euler applies), 4 = second stage of the
// leap-frog step (RHS on T, update S in place)
template <bool MIXED, bool XSW, int phase, int BB>
__global__ void __launch_bounds__(kTPB) k_passB(DevState d) {
  if (blockIdx.x >= (unsigned)d.nblk) return;   // padding block (no partials)
  // neighbours per batch (a tuning parameter: the results do not depend on it)
  constexpr int kB = BB;
  // The loads go out in two rounds, each issued whole before any of its
  // values is used: (1) everything indexed by the list position, including
  // the first kB neighbour-list entries (read whole: the entries past the
  // count are never dereferenced), with the instance records -- the block
  // entry carries the current state set, so the particle's loads need not
  // wait for them; (2) everything the cell, the instance and those entries
  // index.  A wave so waits for two memory round trips before its first
  // neighbour arithmetic.
  const BlockEntry be = d.blk[blockIdx.x];
  const int k = be.inst;
  const bool ok = (int)threadIdx.x < be.nloc;
  const int self = be.pos + (ok ? (int)threadIdx.x : 0);   // padding lanes: a real slot
  const int p = self;
  const int buf = (phase == 1 || phase == 4) ? 1 : 0;
  // the evaluated state, by slot (= list position)
  const int cs = be.cs;
  const double *X = d.x[2 * cs + buf], *Y = d.y[2 * cs + buf], *HS = d.hsml[2 * cs + buf], *RS = d.rho[2 * cs + buf];
  const double *VX = d.vx[2 * cs + buf], *VY = d.vy[2 * cs + buf];
  // (evolving masses: the half-step state's own)
  const double *MS = (XSW && buf && be.tmass) ? d.masst[cs] : d.mass[cs];
  // FP64: pass A's stress; mixed: its per-slot record (d.srec)
  const double *S11 = (const double *)d.sig11, *S12 = (const double *)d.sig12,
               *S22 = (const double *)d.sig22;
  const size_t N = (size_t)d.ntot;
  const int c = d.cellid[cs][self];
  const double xi = X[self], yi = Y[self], hi = HS[self];
  const double rhoi = RS[self], mi = MS[self];
  const double vxi = VX[self], vyi = VY[self];
  const int nn = d.nnb[self];
  const double divi = ((const double *)d.div)[self];
  double s11d = 0., s12d = 0., s22d = 0.;
  float4 reci = make_float4(0.f, 0.f, 0.f, 0.f);
  if (MIXED) {
    reci = d.srec[self];
  } else {
    s11d = S11[self]; s12d = S12[self]; s22d = S22[self];
  }
  int nx[kB];
#pragma unroll
  for (int u = 0; u < kB; ++u) nx[u] = d.nbr[u * N + self];
  const double hthi = d.h[2 * cs + buf][p], Ai = d.A[2 * cs + buf][p];
  double x1 = 0., y1 = 0., h1 = 0., A1 = 0., u1 = 0., v1 = 0.;
  if (phase == 1 || phase == 4) {
    x1 = d.x[2 * cs + 0][p]; y1 = d.y[2 * cs + 0][p];
    h1 = d.h[2 * cs + 0][p]; A1 = d.A[2 * cs + 0][p];
    u1 = d.vx[2 * cs + 0][p]; v1 = d.vy[2 * cs + 0][p];
  }
  const InstDyn &y = d.dyn[k];
  const InstMeta &m = d.meta[k];
  const bool live = ok && !inst_off(y);
  double hsml_new = 1.0, px_new = 0.0, py_new = 0.0;   // for the partials
  double speed = 0.0, dhs = 0.0;
  double rho_new = 1.0, mfo = 0.0, mfc = 0.0;          // Monaghan time step terms
  int st_cnt = 0, st_own = 0;                          // interaction statistics
  if (live) {
    const int W = y.W;
    // third density call of the step (time_integration.f90:262): the
    // corrector's sum on the half-step kernel values with the densities
    // the second call produced (pass A's dsum)
    // (FP64 in both precision modes, as in pass A)
    const bool sumd3 = XSW && phase == 1 && m.den_alg == 1;
    // continuity density in the corrector (time_integration.f90:245-248,
    // 265-268): drhodt on the half-step velocities and gradients, summed
    // with the stress (FP64 in both precision modes)
    const bool cd = XSW && phase == 1 && m.den_alg == 2;
    // summation-density instance: pass A's density (predictor) or its
    // numerator and density (corrector)
    const bool sd = XSW && m.den_alg == 1;
    // regularization (single_step.f90:99-101): the XSPH sum over the pairs,
    // and the artificial viscosity's pair terms after the drag (FP64 in both
    // precision modes)
    const bool xs_on = XSW && m.xsph;
    const bool av_on = XSW && m.avis != 0;
    // smoothing length at the end of the step: Marquis_hsml here (sle = 2),
    // kept (0), or Benz's rate in the end pass (1)
    const bool marq = !XSW || m.sle == 2;
    // the end pass takes the new hsml, the Monaghan terms and the partials
    const bool endp = XSW && m.endp && (phase == 1 || phase == 4);
    double dsi = 1.0, dnumi = 0.;
    if (sd && phase != 2 && phase != 3) dsi = d.dsum[self];
    if (sumd3) dnumi = d.dnum[self];
    // round 2: also the wind at the particle and the cell's run of boundary
    // entries, which the stores below would otherwise hold back
    const int bfl = d.bflag[m.c_off + c];
    double wu, wv;
    wind_at(d, m, xi, yi, wu, wv);
    const int2 run = m.nb > 0 ? d.brun[m.c_off + c] : make_int2(0, 0);
    double kzi = 0.;
    if (sumd3) kzi = d.kz[m.p_off + d.orig[cs][self]];

    // Everything of the update but the velocities is known before the
    // stress sums (single_step.f90:104-106 tendencies; predictor
    // time_integration.f90:207-218, corrector :236-280 on the step-start
    // state x1 ... A1) and is stored now: the registers stay free.
    // (mixed: one division)
    const double dri = MIXED ? divi / rhoi : 0.0;
    const double dhdt = MIXED ? -(hthi * dri) : -(hthi * divi / rhoi);
    const double dAdt = MIXED ? -(Ai * dri) : -(Ai * divi / rhoi);
    const double dt = y.dt;
    if (phase == 0) {
      hsml_new = hi;
      if (!xs_on) {
        px_new = xi + vxi * dt / 2;
        py_new = yi + vyi * dt / 2;
        d.x[2 * cs + 1][p] = px_new;
        d.y[2 * cs + 1][p] = py_new;
      }
      const double hh = hthi + dhdt * dt / 2;
      double Ah = Ai + dAdt * dt / 2;
      if (Ah > 1) Ah = 1;
      d.h[2 * cs + 1][p] = hh;
      d.A[2 * cs + 1][p] = Ah;
      // density (time_integration.f90:218): diagnostic, or the summation
      // pass A computed on the step-start kernel values
      d.rho[2 * cs + 1][p] = sd ? dsi : hh * kRhoIce;
      d.hsml[2 * cs + 1][p] = hi;
      // mass_evolution (:227, :352) of evolving masses
      if (XSW && be.tmass) d.masst[cs][p] = area_mass(Ah, hh, hi);
    } else if (phase == 1) {
      const double h12 = h1 + dhdt * dt / 2, A12 = A1 + dAdt * dt / 2;
      if (!xs_on) {
        const double x12 = x1 + vxi * dt / 2, y12 = y1 + vyi * dt / 2;
        px_new = 2 * x12 - x1;
        py_new = 2 * y12 - y1;
        d.x[2 * cs + 0][p] = px_new;
        d.y[2 * cs + 0][p] = py_new;
      }
      const double hn = 2 * h12 - h1;
      double An = 2 * A12 - A1;
      if (An > 1) An = 1;
      d.h[2 * cs + 0][p] = hn;
      d.A[2 * cs + 0][p] = An;
      hsml_new = hi;
      // mass_evolution (:271), before smoothing_length
      if (XSW && be.tmass) d.mass[cs][p] = area_mass(An, hn, hi);
      if (!sumd3 && !cd) {
        const double rn = hn * kRhoIce;
        d.rho[2 * cs + 0][p] = rn;
        rho_new = rn;
        if (marq) {
          // Marquis_hsml (smoothing_length.f90:136-142)
          double hs = sqrt(mi / rn) * kHsmlMult;
          if (hs > m.max_hsml) hs = m.max_hsml;
          d.hsml[2 * cs + 0][p] = hs;
          hsml_new = hs;
        }
      }
    } else if (phase == 4) {
      // leapfrog_scheme (time_integration.f90:361-391): thickness,
      // position and concentration from the step-start state and this
      // evaluation's tendencies (dxdt = the half-step velocity), density on
      // this evaluation (diagnostic, or pass A's summation), Marquis_hsml
      const double hn = h1 + dhdt * dt;
      if (!xs_on) {
        px_new = x1 + (u1 + vxi) * dt / 2;
        py_new = y1 + (v1 + vyi) * dt / 2;
        d.x[2 * cs + 0][p] = px_new;
        d.y[2 * cs + 0][p] = py_new;
      }
      double An = A1 + dAdt * dt;
      if (An > 1) An = 1;
      d.h[2 * cs + 0][p] = hn;
      d.A[2 * cs + 0][p] = An;
      hsml_new = hi;
      // mass_evolution (:380), before smoothing_length
      if (XSW && be.tmass) d.mass[cs][p] = area_mass(An, hn, hi);
      // (continuity density: k_condens takes the density, hsml and the
      // partials from here, on the new velocities)
      if (!(XSW && m.den_alg == 2)) {
        const double rn = sd ? dsi : hn * kRhoIce;
        d.rho[2 * cs + 0][p] = rn;
        rho_new = rn;
        if (marq) {
          double hs = sqrt(mi / rn) * kHsmlMult;
          if (hs > m.max_hsml) hs = m.max_hsml;
          d.hsml[2 * cs + 0][p] = hs;
          hsml_new = hs;
        }
      }
    }
    const bool listed = nn <= kNbrMax;
    int sj[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) sj[u] = (listed && u < nn) ? nx[u] : self;

    double ax = 0., ay = 0.;
    int nbnd = 0;          // boundary pairs of this particle
    // stress_1 (internal_stress.f90:103-136), over pass A's neighbour list
    // in batches of kB neighbours whose loads are issued together
    // waves with many or boundary neighbours get issue priority
    if (__any(nn > 10 || (m.nb > 0 && bfl == y.lstamp))) __builtin_amdgcn_s_setprio(2);
    const double ri2 = rhoi * rhoi;
    const float s11f = reci.x, s12f = reci.y, s22f = reci.z;
    KSum sx, sy;
    double den3 = 0.;
    if (sumd3) den3 = kzi * mi / dsi;
    const int skf = XSW ? m.skf : 6;
    const double ksc = XSW ? kern_scale(skf) : 1.0;
    double drho = 0.;
    auto cden_pair = [&](double dx, double dy, double mh, double mq, double uq, double vq) {
      const double r = sqrt(dx * dx + dy * dy);
      const double gg = kern_grad(skf, r, mh);
      const double dw0 = gg * (dx / r) * ksc, dw1 = gg * (dy / r) * ksc;
      const double w1 = (vxi - uq) * dw0, w4 = (vyi - vq) * dw1;
      drho = drho + mq * (w1 + w4);
    };
    auto dens_pair = [&](double dx, double dy, double mh, double mq, double rho2) {
      const double w = kern_w(skf, sqrt(dx * dx + dy * dy), mh);
      den3 = den3 + mq * w / rho2;
    };
    // XSPH (regularization.f90:73-91): the pair's dvx = v_i - v_j enters both
    // particles' sums with the same sign; owned pairs (self = i) are the
    // tail of the list, nI + nX entries
    double xsx = 0., xsy = 0.;
    const int own0 = xs_on ? nn - (d.own_cnt[self] & 1023) - ((d.own_cnt[self] >> 10) & 1023) : 0;
    auto xsph_pair = [&](bool own, double dx, double dy, double mh, double mq, double rq,
                         double uq, double vq) {
      const double w = kern_w(skf, sqrt(dx * dx + dy * dy), mh);
      const double rm = 0.5 * (rhoi + rq);
      const double e0 = own ? vxi - uq : uq - vxi;
      const double e1 = own ? vyi - vq : vq - vyi;
      xsx = xsx + mq * e0 * w / rm;
      xsy = xsy + mq * e1 * w / rm;
    };
    const bool pa2 = XSW && m.pa_sph == 2;
    auto stress_pair = [&](double dx, double dy, double r, double mh, double rq, double mq,
                           double q11, double q12, double q22) {
      const double gg = kern_grad(skf, r, mh);
      const double dw0 = gg * (dx / r) * ksc, dw1 = gg * (dy / r) * ksc;
      if (pa2) {
        // stress_2 (internal_stress.f90:189-216); as the pair's second
        // particle the reference subtracts the term of the opposite gradient,
        // the same value
        const double coef = mq / kRhoIce / hthi / rq;
        double cj = q11 * dw0 + q12 * dw1;
        double ci = s11d * dw0 + s12d * dw1;
        ax = ax + coef * (ci + cj);
        cj = q12 * dw0 + q22 * dw1;
        ci = s12d * dw0 + s22d * dw1;
        ay = ay + coef * (ci + cj);
        return;
      }
      const double rq2 = rq * rq;
      const double coef = rhoi * mq / kRhoIce / hthi;
      double cj = (q11 * dw0 + q12 * dw1) / rq2;
      double ci = (s11d * dw0 + s12d * dw1) / ri2;
      ax = ax + coef * (ci + cj);
      cj = (q12 * dw0 + q22 * dw1) / rq2;
      ci = (s12d * dw0 + s22d * dw1) / ri2;
      ay = ay + coef * (ci + cj);
    };
    // mixed: q = the neighbour's record {sig / rho^2, mass}
    auto stress_pair_f = [&](double dx, double dy, double mh, float4 q) {
      const float fx = (float)dx, fy = (float)dy;
      const float gr = kern_gradr_f(skf, fx * fx + fy * fy, (float)mh);
      const float dw0 = fx * gr, dw1 = fy * gr;
      const float t11 = q.x + s11f, t12 = q.y + s12f, t22 = q.z + s22f;
      sx.add(q.w * (t11 * dw0 + t12 * dw1));
      sy.add(q.w * (t12 * dw0 + t22 * dw1));
    };
    if (listed) {
      for (int j0 = 0; j0 < nn; j0 += kB) {
        double xq[kB], yq[kB], hq[kB], rq[kB], mq[kB], r2[kB];
        double q11[kB], q12[kB], q22[kB], uq[kB], vq[kB];
        float4 qr[kB];
#pragma unroll
        for (int u = 0; u < kB; ++u) {
          const int t = sj[u];
          xq[u] = X[t]; yq[u] = Y[t]; hq[u] = HS[t];
          uq[u] = (cd || xs_on) ? VX[t] : 0.0;
          vq[u] = (cd || xs_on) ? VY[t] : 0.0;
          if (MIXED) {
            qr[u] = d.srec[t];
            mq[u] = (sumd3 || cd || xs_on) ? MS[t] : 0.0;
            rq[u] = xs_on ? RS[t] : 1.0;
          } else {
            rq[u] = RS[t]; mq[u] = MS[t];
            q11[u] = S11[t]; q12[u] = S12[t]; q22[u] = S22[t];
          }
          r2[u] = sumd3 ? d.dsum[t] : 1.0;
        }
        // the next batch's list entries go out with this batch's data
        // (past the end of the list, the last entry again: no line of the
        // list is read that the particle does not use)
        const int j1 = j0 + kB;
#pragma unroll
        for (int u = 0; u < kB; ++u) nx[u] = d.nbr[min(j1 + u, nn - 1) * N + self];
#pragma unroll
        for (int u = 0; u < kB; ++u) {
          // (mixed: continue keeps the batch's loads issued together; the
          // FP64 variant, short of registers, schedules better with break)
          if (j0 + u >= nn) {
            if (MIXED) continue;
            break;
          }
          const double dx = xi - xq[u], dy = yi - yq[u];
          const double mh = (hi + hq[u]) / 2.;
          if (MIXED)
            stress_pair_f(dx, dy, mh, qr[u]);
          else
            stress_pair(dx, dy, sqrt(dx * dx + dy * dy), mh, rq[u], mq[u], q11[u], q12[u],
                        q22[u]);
          if (sumd3) dens_pair(dx, dy, mh, mq[u], r2[u]);
          if (cd) cden_pair(dx, dy, mh, mq[u], uq[u], vq[u]);
          if (xs_on) xsph_pair(j0 + u >= own0, dx, dy, mh, mq[u], rq[u], uq[u], vq[u]);
        }
#pragma unroll
        for (int u = 0; u < kB; ++u) sj[u] = j1 + u < nn ? nx[u] : self;
      }
    } else {
      const int cy = c / W, cx = c - cy * W;
      const Ranges R = neighbour_ranges(d.cell_start + m.c_off, W, cx, cy);
      for (int g = 0; g < 6; ++g) {
        for (int s = R.a[g]; s < R.b[g]; ++s) {
          if (s == self) continue;
          const double dx = xi - X[s], dy = yi - Y[s];
          const double mh = (hi + HS[s]) / 2.;
          double r;
          if (!in_support(dx * dx + dy * dy, mh, r)) continue;
          if (MIXED)
            stress_pair_f(dx, dy, mh, d.srec[s]);
          else
            stress_pair(dx, dy, r, mh, RS[s], MS[s], S11[s], S12[s], S22[s]);
          if (sumd3) dens_pair(dx, dy, mh, MS[s], d.dsum[s]);
          if (cd) cden_pair(dx, dy, mh, MS[s], VX[s], VY[s]);
          if (xs_on) xsph_pair(pair_kind(R, s, self) != 0, dx, dy, mh, MS[s], RS[s], VX[s], VY[s]);
        }
      }
    }
    if (MIXED) {
      const double cf = (pa2 ? 1.0 : rhoi) / kRhoIce / hthi;
      ax = cf * sx.value();
      ay = cf * sy.value();
    }

    // boundary_kernel -> b_Wendland_C6 -> bf_Monaghan, boundary pairs in
    // ascending boundary index (boundary_kernel.f90:477-502)
    if (m.nb > 0 && bfl == y.lstamp) {
      // The cell's boundary particles in ascending index, each with the
      // slot this cell takes in its 3x3 block.  A pair's rank counts the
      // cell's particles before this one (list order) in range of the same
      // boundary particle.  The lanes of one cell share the cell's entry
      // run, so the entries are walked wave-synchronously (entry u of every
      // lane's run at once): the in-range flags of the cell's lanes are one
      // ballot, and the count over the cell's earlier lanes is a popcount;
      // only the members in an earlier wave are tested one by one.  The
      // loads of kBU entries go out together.
      constexpr int kBU = 2;
      const int own0 = d.cell_start[m.c_off + c];   // first member of the own cell
      const int lane = (int)threadIdx.x & 63;
      const int wave0 = self - lane;                 // slot of the wave's lane 0
      const int first = max(own0 - wave0, 0);        // the cell's first lane here
      const unsigned long long below =
          (((unsigned long long)1 << lane) - 1) & ~(((unsigned long long)1 << first) - 1);
      int nbn = 0;
      float fbx = 0.0f, fby = 0.0f;   // mixed: the forces' sum
      const bool rfresh = phase != 1 && y.rfresh;
      const int bfa = XSW ? m.bfa : 1;
      const bool bpres = !XSW || m.bpres, fric = XSW && m.fric;
      const int *bl = d.blist + 9 * (size_t)m.b_off;
      for (int u0 = 0; __any(run.x + u0 < run.y); u0 += kBU) {
        int ent[kBU], rb[kBU];
        double bx[kBU], by[kBU], mb[kBU], hk[kBU];
        bool inr[kBU];
#pragma unroll
        for (int j = 0; j < kBU; ++j) {
          const int e = run.x + u0 + j;
          ent[j] = e < run.y ? bl[e] : -1;
        }
#pragma unroll
        for (int j = 0; j < kBU; ++j) {
          const int bg = m.b_off + (ent[j] < 0 ? 0 : (ent[j] & 0xffffff));
          bx[j] = d.xb[bg];
          by[j] = d.yb[bg];
          mb[j] = d.massb[bg];
          rb[j] = d.bbase[bg] + d.bpre[bg * 9 + (ent[j] < 0 ? 0 : (ent[j] >> 24))];
        }
#pragma unroll
        for (int j = 0; j < kBU; ++j) {
          const double bdx = bx[j] - xi, bdy = by[j] - yi;
          inr[j] = ent[j] >= 0 && sqrt(bdx * bdx + bdy * bdy) <= kBLength;
          int before = __popcll(__ballot(inr[j]) & below);
          if (inr[j]) {
            for (int s = own0; s < wave0; ++s) {
              const double ex = bx[j] - X[s], ey = by[j] - Y[s];
              if (sqrt(ex * ex + ey * ey) <= kBLength) ++before;
            }
          }
          const int rank = rb[j] + before;
          hk[j] = 0.0;
          if (inr[j] && rank < m.q_cap) {
            // the entry of a pair below the rank limit is refreshed from the
            // pair's smoothing lengths (and kept for later stale reads)
            const int2 hp = rfresh ? d.hpair[m.q_off + rank] : make_int2(-1, -1);
            if (hp.x >= 0) {
              hk[j] = (HS[hp.x] + HS[hp.y]) / 2.;
              d.hq[m.q_off + rank] = hk[j];
            } else {
              hk[j] = d.hq[m.q_off + rank];
            }
          }
        }
#pragma unroll
        for (int j = 0; j < kBU; ++j) {
          if (!inr[j]) continue;
          ++nbn;
          const double bdx = bx[j] - xi, bdy = by[j] - yi;
          if (bfa == 3) {
            // bf_Marquis (:210-221): no force with the boundary pressure on;
            // the friction brakes the particle, less with distance (FP64)
            if (bpres || !fric) continue;
            const double br = sqrt(bdx * bdx + bdy * bdy);
            ax = ax - vxi * (1 - (br / kBLength));
            ay = ay - vyi * (1 - (br / kBLength));
            continue;
          }
          if (bfa == 2) {
            // bf_Wang (boundary_force.f90:145-174): the normal velocity and
            // the rest of the speed, each as a momentum over dt; friction
            // caps the tangential part at xi times the normal (FP64)
            if (!bpres) continue;
            const int bg = m.b_off + (ent[j] & 0xffffff);
            const double n0 = d.bnx[bg], n1 = d.bny[bg];
            const double v0 = n0 * vxi + n1 * vyi;
            double v1 = sqrt(vxi * vxi + vyi * vyi) - v0;
            const double F0 = mi * v0 / dt;
            double F1 = mi * v1 / dt;
            if (fric && fabs(v1) > fabs(kXi * v0)) {
              v1 = v1 * kXi * fabs(F0 / F1);
              F1 = mi * v1 / dt;
            }
            ax = ax + (F0 * n0 + F1 * n1);
            ay = ay + (F0 * n1 - F1 * n0);
            continue;
          }
          if (fric) {
            // bf_Monaghan's normal force (FP64 here), then its friction
            // (:87-100): kappa_t c v . bdx / br / rho against the boundary
            // kernel's gradient
            const double br = sqrt(bdx * bdx + bdy * bdy);
            if (bpres) {
              const double bw = bkern_w(skf, br, hk[j]);
              const double den = (mi + mb[j]) * (br * br);
              const double F0 = kKappaN * bdx * bw * 2 * mb[j] / den;
              const double F1 = kKappaN * bdy * bw * 2 * mb[j] / den;
              ax = ax + (-F0) / rhoi;
              ay = ay + (-F1) / rhoi;
            }
            const double g = bkern_g(skf, br, hk[j]);
            const double bw0 = -(g * (bdx / br) * ksc), bw1 = -(g * (bdy / br) * ksc);
            const double v_r = vxi * bdx + vyi * bdy;
            const double pi_aj = kKappaT * sound_speed(rhoi) * v_r / br / rhoi;
            ax = ax + (-(mb[j] * pi_aj * bw0)) / rhoi;
            ay = ay + (-(mb[j] * pi_aj * bw1)) / rhoi;
            continue;
          }
          if (!bpres) continue;
          if (MIXED) {
            // mixed: the force in FP32 (hardware sqrt and reciprocals);
            // the pair set and its ranks stay those of the FP64 test
            const float fx = (float)bdx, fy = (float)bdy;
            const float br2 = fx * fx + fy * fy;
            const float q = __builtin_amdgcn_sqrtf(br2) * (float)(1.0 / kBLength);
            const float bw = kern_wendland(skf) ? bkern_w_f(skf, q, (float)hk[j])
                                                : (float)bkern_w(skf, sqrt(bdx * bdx + bdy * bdy),
                                                                 hk[j]);
            const float cf = (float)(2.0 * kKappaN) * bw * (float)mb[j] *
                             __builtin_amdgcn_rcpf((float)(mi + mb[j]) * br2);
            fbx += cf * fx;
            fby += cf * fy;
            continue;
          }
          const double br = sqrt(bdx * bdx + bdy * bdy);
          const double bw = bkern_w(skf, br, hk[j]);
          const double den = (mi + mb[j]) * (br * br);
          const double F0 = kKappaN * bdx * bw * 2 * mb[j] / den;
          const double F1 = kKappaN * bdy * bw * 2 * mb[j] / den;
          ax = ax + (-F0) / rhoi;
          ay = ay + (-F1) / rhoi;
        }
      }
      if (MIXED) {
        ax = ax - (double)fbx / rhoi;
        ay = ay - (double)fby / rhoi;
      }
      nbnd = nbn;
    }

    // interaction_stats (nnps.f90:613-668) at the steps the reference runs
    // it, and in every single evaluation: the capacity STOP, the statistics
    if (phase == 2 || y.itimestep % kPrintStep == 0) {
      const int oc = d.own_cnt[self];
      st_cnt = (oc >> 20) + nbnd;
      st_own = (oc & 1023) + ((oc >> 10) & 1023);
      if (y.itimestep % kPrintStep == 0 && st_cnt > kMaxIndInt)
        atomicOr(&d.dyn[k].err, kErrCapacity);
    }

    // wind_forcing / water_drag (wind_forcing.f90:35-50, water_drag.f90:35-51)
    // (each gated by its scenario switch, globals.f90:182-183)
    const bool won = !XSW || m.wind_on, con = !XSW || m.water_on;
    if (y.time > 0.) {
      const double du = wu - vxi, dv = wv - vyi;
      const double hr = hthi * kRhoIce;
      const double cu = m.cur_u - vxi, cv = m.cur_v - vyi;
      if (MIXED) {
        // one division for the four terms
        const double ihr = 1.0 / hr;
        const double wx = won ? kCWind * kRhoAir * du * fabs(du) : 0.0;
        const double wy = won ? kCWind * kRhoAir * dv * fabs(dv) : 0.0;
        const double cx = con ? kCWater * kRhoWater * cu * fabs(cu) : 0.0;
        const double cy = con ? kCWater * kRhoWater * cv * fabs(cv) : 0.0;
        ax = ax + (wx + cx) * ihr;
        ay = ay + (wy + cy) * ihr;
      } else {
        if (won) {
          ax = ax + (kCWind * kRhoAir * du * fabs(du)) / hr;
          ay = ay + (kCWind * kRhoAir * dv * fabs(dv)) / hr;
        }
        if (con) {
          ax = ax + (kCWater * kRhoWater * cu * fabs(cu)) / hr;
          ay = ay + (kCWater * kRhoWater * cv * fabs(cv)) / hr;
        }
      }
    }
    // the artificial viscosity's pair terms (regularization.f90:198-208,
    // :293-303) in list order, with pi_ij of k_avpair / k_avres; as the
    // pair's second particle the reference adds the term of the opposite
    // gradient, the same value
    if (av_on && listed) {
      for (int j = 0; j < nn; ++j) {
        const int t = d.nbr[j * N + self];
        const double pi = d.piv[d.pidx[j * N + self]];
        const double dx = xi - X[t], dy = yi - Y[t];
        const double r = sqrt(dx * dx + dy * dy);
        const double gg = kern_grad(skf, r, (hi + HS[t]) / 2.);
        const double dw0 = gg * (dx / r) * ksc, dw1 = gg * (dy / r) * ksc;
        ax = ax - MS[t] * pi * dw0;
        ay = ay - MS[t] * pi * dw1;
      }
    }
    // XSPH: dxdt = 0 - chi xsph_sum (regularization.f90:93), then + vx
    // (single_step.f90:106); the position updates deferred above
    double dxdt0 = vxi, dxdt1 = vyi;
    if (xs_on) {
      dxdt0 = (0.0 - kChi * xsx) + vxi;
      dxdt1 = (0.0 - kChi * xsy) + vyi;
      if (phase == 0) {
        px_new = xi + dxdt0 * dt / 2;
        py_new = yi + dxdt1 * dt / 2;
        d.x[2 * cs + 1][p] = px_new;
        d.y[2 * cs + 1][p] = py_new;
      } else if (phase == 1) {
        const double x12 = x1 + dxdt0 * dt / 2, y12 = y1 + dxdt1 * dt / 2;
        px_new = 2 * x12 - x1;
        py_new = 2 * y12 - y1;
        d.x[2 * cs + 0][p] = px_new;
        d.y[2 * cs + 0][p] = py_new;
      } else if (phase == 4) {
        px_new = x1 + (u1 + dxdt0) * dt / 2;
        py_new = y1 + (v1 + dxdt1) * dt / 2;
        d.x[2 * cs + 0][p] = px_new;
        d.y[2 * cs + 0][p] = py_new;
      }
    }
    // velocities (time_integration.f90:210-211, 240-241) or tendencies
    if (phase == 0) {
      d.vx[2 * cs + 1][p] = vxi + ax * dt / 2;
      d.vy[2 * cs + 1][p] = vyi + ay * dt / 2;
    } else if (phase == 1) {
      const double u12 = u1 + ax * dt / 2, v12 = v1 + ay * dt / 2;
      d.vx[2 * cs + 0][p] = 2 * u12 - u1;
      d.vy[2 * cs + 0][p] = 2 * v12 - v1;
      if (sumd3 || cd) {
        double rn;
        if (sumd3) {
          // summation density of the corrector
          rn = dnumi / den3;
        } else {
          // rho_12 = rho_1 + drhodt dt/2, rho = 2 rho_12 - rho_1
          const double r1 = d.rho[2 * cs + 0][p];
          const double r12 = r1 + drho * dt / 2;
          rn = 2 * r12 - r1;
        }
        d.rho[2 * cs + 0][p] = rn;
        rho_new = rn;
        if (marq) {
          // Marquis_hsml
          double hs = sqrt(mi / rn) * kHsmlMult;
          if (hs > m.max_hsml) hs = m.max_hsml;
          d.hsml[2 * cs + 0][p] = hs;
          hsml_new = hs;
        }
      }
      if (endp) {
        // for the end pass's Monaghan time step terms
        d.tend[2 * (size_t)d.ntot + p] = ax;
        d.tend[3 * (size_t)d.ntot + p] = ay;
      }
    } else if (phase == 4) {
      d.vx[2 * cs + 0][p] = u1 + ax * dt;
      d.vy[2 * cs + 0][p] = v1 + ay * dt;
      if (XSW && (m.den_alg == 2 || endp)) {
        // for k_condens's or the end pass's Monaghan time step terms
        d.tend[2 * (size_t)d.ntot + p] = ax;
        d.tend[3 * (size_t)d.ntot + p] = ay;
      }
    } else {
      const int N = d.ntot;
      d.tend[p] = dxdt0;
      d.tend[N + p] = dxdt1;
      d.tend[2 * N + p] = ax;
      d.tend[3 * N + p] = ay;
      d.tend[4 * N + p] = dhdt;
      d.tend[5 * N + p] = dAdt;
    }
    // the candidate lists' movement bound: speed of the evaluated state and
    // hsml change of the new one (final only here: the summation and
    // continuity densities of the corrector set it with the velocities)
    speed = sqrt(vxi * vxi + vyi * vyi);
    dhs = fabs(hsml_new - hi);
    // time_step at the end of the step (time_integration.f90:280, :389):
    // the Monaghan step's terms from this evaluation's dvxdt
    if (XSW && (phase == 1 || phase == 4) && m.tse == 1)
      monaghan_terms(hsml_new, rho_new, ax, ay, mfo, mfc);
  }
  // (with every lane of the wave: a wave holds one instance)
  if (phase == 2 || y.itimestep % kPrintStep == 0) stats_add(d.dyn[k], live, st_cnt, st_own);
  if constexpr (XSW && (phase == 1 || phase == 4)) {
    // (block-uniform: k_condens writes a leap-frog continuity-density
    // instance's partials)
    if (!(phase == 4 && m.den_alg == 2) && !m.endp)
      block_partials<kParts>(d, k, live, px_new, py_new, hsml_new, speed, dhs, mfo, mfc);
  }
  else if constexpr (phase != 2 && phase != 3)
    block_partials(d, k, live, px_new, py_new, hsml_new, speed, dhs);
}

// euler_scheme's update (time_integration.f90:117-146) after the Euler
// evaluation: the state advanced in place by dt with the tendencies pass B
// wrote, the density of the new state (diagnostic, or pass A's summation
// on this evaluation's kernel values and densities), Marquis_hsml, and the
// block's partials for the next grid.  Nothing here reads a neighbour, so
// the update is in place.
__global__ void __launch_bounds__(kTPB) k_euler(DevState d) {
  if (blockIdx.x >= (unsigned)d.nblk) return;
  int p = 0, k = 0;
  const bool ok = particle_of(d, p, k);
  const InstDyn &y = d.dyn[k];
  const bool live = ok && !inst_off(y);
  double xn = 0., yn = 0., hs = 1.0, sp = 0., dh = 0., fo = 0., fc = 0.;
  if (live) {
    const InstMeta &m = d.meta[k];
    const int cs = d.blk[blockIdx.x].cs;
    const size_t N = (size_t)d.ntot;
    const double *T = d.tend;
    const double dt = y.dt;
    const double vx = d.vx[2 * cs][p], vy = d.vy[2 * cs][p];
    xn = d.x[2 * cs][p] + T[p] * dt;
    yn = d.y[2 * cs][p] + T[N + p] * dt;
    const double un = vx + T[2 * N + p] * dt, vn = vy + T[3 * N + p] * dt;
    const double hn = d.h[2 * cs][p] + T[4 * N + p] * dt;
    double An = d.A[2 * cs][p] + T[5 * N + p] * dt;
    if (An > 1) An = 1;
    // (summation density: pass A's sum; continuity density: k_condens's)
    const double rn = m.den_alg != 3 ? d.dsum[p] : hn * kRhoIce;
    const double h0 = d.hsml[2 * cs][p];
    if (m.endp) {
      hs = d.hnew[p];                 // the end pass's (Benz's or Marquis')
    } else if (m.sle == 2) {
      hs = sqrt(d.mass[cs][p] / rn) * kHsmlMult;
      if (hs > m.max_hsml) hs = m.max_hsml;
    } else {
      hs = h0;                        // sle = 0: kept
    }
    sp = sqrt(vx * vx + vy * vy);
    dh = fabs(hs - h0);
    // time_step (time_integration.f90:146): the Monaghan step's terms
    if (m.tse == 1) monaghan_terms(hs, rn, T[2 * N + p], T[3 * N + p], fo, fc);
    if (m.avdt) fc = d.cfl[p];
    d.x[2 * cs][p] = xn;
    d.y[2 * cs][p] = yn;
    d.vx[2 * cs][p] = un;
    d.vy[2 * cs][p] = vn;
    d.h[2 * cs][p] = hn;
    d.A[2 * cs][p] = An;
    d.rho[2 * cs][p] = rn;
    d.hsml[2 * cs][p] = hs;
    // mass_evolution (:137), with the hsml before smoothing_length
    if (m.sle != 2) d.mass[cs][p] = area_mass(An, hn, h0);
  }
  block_partials<kParts>(d, k, live, xn, yn, hs, sp, dh, fo, fc);
}

// con_density (density.f90:107-165) after an update, for the instances with
// den_alg = 2: drhodt = sum over the evaluation's pairs, in the order the
// reference's list reaches the particle, of m_j (v_i - v_j) . dwdx_ij, with
// the evaluation's gradients and the velocities the update produced (FP64 in
// both precision modes; the product does not depend on which particle of the
// pair comes first).
//   MODE 0, after the predictor (time_integration.f90:221-224, :347-349):
//     gradients at the step-start state, half-step velocities;
//     rho_T = rho_S + drhodt dt / 2.
//   MODE 1, after the Euler evaluation and before its in-place update
//     (:131-134): velocities v + dvxdt dt from the evaluation's tendencies
//     (the update's own expression); dsum = rho + drhodt dt for k_euler.
//   MODE 2, after the leap-frog's second evaluation (:375-377): gradients at
//     the half-step state, the new velocities; rho = rho_1 + drhodt dt, then
//     Marquis_hsml and the block's partials (which pass B leaves to this
//     kernel for such instances).
template <int MODE>
__global__ void __launch_bounds__(kTPB) k_condens(DevState d) {
  if (blockIdx.x >= (unsigned)d.nblk) return;
  const BlockEntry be = d.blk[blockIdx.x];
  const int k = be.inst;
  const InstMeta &m = d.meta[k];
  if (m.den_alg != 2) return;   // (block-uniform)
  const InstDyn &y = d.dyn[k];
  const bool ok = (int)threadIdx.x < be.nloc;
  const bool live = ok && !inst_off(y);
  const int self = be.pos + (ok ? (int)threadIdx.x : 0);
  const int cs = be.cs;
  const int eb = MODE == 2 ? 1 : 0;   // buffer of the evaluated state
  const int vb = MODE == 0 ? 1 : 0;   // buffer of the velocities
  const double *X = d.x[2 * cs + eb], *Y = d.y[2 * cs + eb], *HS = d.hsml[2 * cs + eb];
  const double *VX = d.vx[2 * cs + vb], *VY = d.vy[2 * cs + vb];
  // (evolving masses: after the leap-frog's second evaluation the density
  // call precedes mass_evolution, time_integration.f90:372-380)
  const double *MS = (MODE == 2 && be.tmass) ? d.masst[cs] : d.mass[cs];
  const size_t N = (size_t)d.ntot;
  const double dt = y.dt;
  double px = 0., py = 0., hs = 1.0, sp = 0., dh = 0., fo = 0., fc = 0.;
  if (live) {
    // the velocity of slot t after the update
    auto vel = [&](int t, double &u, double &v) {
      u = VX[t];
      v = VY[t];
      if (MODE == 1) {
        u = u + d.tend[2 * N + t] * dt;
        v = v + d.tend[3 * N + t] * dt;
      }
    };
    const int skf = m.skf;
    const double ksc = kern_scale(skf);
    const double xi = X[self], yi = Y[self], hi = HS[self];
    double ui, vi;
    vel(self, ui, vi);
    double drho = 0.;
    auto pair = [&](int t, double dx, double dy, double r, double mh) {
      double uq, vq;
      vel(t, uq, vq);
      const double gg = kern_grad(skf, r, mh);
      const double dw0 = gg * (dx / r) * ksc, dw1 = gg * (dy / r) * ksc;
      const double w1 = (ui - uq) * dw0, w4 = (vi - vq) * dw1;
      drho = drho + MS[t] * (w1 + w4);
    };
    const int nn = d.nnb[self];
    if (nn <= kNbrMax) {
      for (int j = 0; j < nn; ++j) {
        const int t = d.nbr[j * N + self];
        const double dx = xi - X[t], dy = yi - Y[t];
        pair(t, dx, dy, sqrt(dx * dx + dy * dy), (hi + HS[t]) / 2.);
      }
    } else {
      const int W = y.W;
      const int c = d.cellid[cs][self];
      const int cy = c / W, cx = c - cy * W;
      const Ranges R = neighbour_ranges(d.cell_start + m.c_off, W, cx, cy);
      for (int g = 0; g < 6; ++g) {
        for (int t = R.a[g]; t < R.b[g]; ++t) {
          if (t == self) continue;
          const double dx = xi - X[t], dy = yi - Y[t];
          const double mh = (hi + HS[t]) / 2.;
          double r;
          if (!in_support(dx * dx + dy * dy, mh, r)) continue;
          pair(t, dx, dy, r, mh);
        }
      }
    }
    const double rho0 = d.rho[2 * cs + 0][self];
    if (MODE == 0) {
      d.rho[2 * cs + 1][self] = rho0 + drho * dt / 2;
    } else if (MODE == 1) {
      d.dsum[self] = rho0 + drho * dt;
    } else {
      const double rn = rho0 + drho * dt;
      d.rho[2 * cs + 0][self] = rn;
      hs = hi;
      if (m.sle == 2) {
        // Marquis_hsml (smoothing_length.f90:136-142)
        hs = sqrt(MS[self] / rn) * kHsmlMult;
        if (hs > m.max_hsml) hs = m.max_hsml;
        d.hsml[2 * cs + 0][self] = hs;
      }
      px = d.x[2 * cs + 0][self];
      py = d.y[2 * cs + 0][self];
      const double ue = d.vx[2 * cs + 1][self], ve = d.vy[2 * cs + 1][self];
      sp = sqrt(ue * ue + ve * ve);
      dh = fabs(hs - hi);
      if (m.tse == 1) monaghan_terms(hs, rn, d.tend[2 * N + self], d.tend[3 * N + self], fo, fc);
    }
  }
  // (an instance with the end pass gets its partials there)
  if (MODE == 2 && !m.endp) block_partials<kParts>(d, k, live, px, py, hs, sp, dh, fo, fc);
}

// ----------------------------------------- artificial viscosity and end pass

// art_visc_HK and art_visc_Mon (regularization.f90:179-211, :268-306) walk
// the pair list in the reference's order, and pi_ij is not reset per pair:
// a pair whose criterion v_ij . r_ij <= 0 fails keeps the value of the last
// pair before it that set one (0 before the first).  k_avpair gives every
// pair owned by a slot its rank in the packed pair ranks (pair_start, moff
// and cell_intra of k_pairtot and the scan: per cell all intra pairs, then
// all cross pairs, each in member order), whether it sets pi_ij and the
// value it sets, and records the rank on both particles' list entries (the
// owner's entry and the partner's, which lies in the partner's non-owned
// head); k_avlast and k_avres then give every other pair the value of the
// last setting pair of its instance before it.  Pass B adds the pair terms
// in list order.  FP64 in both precision modes.
__global__ void __launch_bounds__(kTPB) k_avpair(DevState d, int buf) {
  int p = 0, k = 0;
  if (!particle_of(d, p, k)) return;
  const InstMeta &m = d.meta[k];
  const InstDyn &y = d.dyn[k];
  if (m.avis == 0 || inst_off(y)) return;
  const int cs = d.blk[blockIdx.x].cs, sb = 2 * cs + buf;
  const double *X = d.x[sb], *Y = d.y[sb], *VX = d.vx[sb], *VY = d.vy[sb];
  const double *HS = d.hsml[sb], *RS = d.rho[sb];
  const size_t N = (size_t)d.ntot;
  const int nn = d.nnb[p];
  if (nn > kNbrMax) {         // the pair terms need the stored lists
    atomicOr(&d.dyn[k].err, kErrCapacity);
    return;
  }
  const int oc = d.own_cnt[p];
  const int nI = oc & 1023, nX = (oc >> 10) & 1023;
  const int c = d.cellid[cs][p];
  const int base = d.pair_start[m.c_off + c];
  const int mo = d.moff[p];
  const int cint = d.cell_intra[m.c_off + c];
  const int j0 = nn - nI - nX, jx = nn - nX;
  const double xi = X[p], yi = Y[p], vxi = VX[p], vyi = VY[p], hi = HS[p], rhoi = RS[p];
  const double ci = m.needc ? sound_speed(rhoi) : 0.0;
  const double lvi = m.avis == 1 ? d.lv[p] : 0.0;
  for (int j = j0; j < nn; ++j) {
    const int q = d.nbr[j * N + p];
    const int g = j < jx ? base + (mo & 0xffff) + (j - j0) : base + cint + (mo >> 16) + (j - jx);
    if (g >= d.pcap) {
      atomicOr(&d.dyn[k].err, kErrCapacity);
      return;
    }
    // pair (i, j) = (p, q): p owns it, so it is the pair's first particle
    const double dx = xi - X[q], dy = yi - Y[q];
    const double dv0 = vxi - VX[q], dv1 = vyi - VY[q];
    const double crit = dv0 * dx + dv1 * dy;
    const double rq = RS[q];
    double pi = 0.0;
    if (m.avis == 1) {
      pi = lvi / (rhoi * rhoi) + d.lv[q] / (rq * rq);
    } else {
      const double mh = (hi + HS[q]) / 2.;
      const double r = sqrt(dx * dx + dy * dy);
      const double s_cor = (double)0.01f * mh;
      const double iota = mh * crit / (r * r + s_cor * s_cor);
      const double cq = m.needc ? sound_speed(rq) : 0.0;
      pi = (kBetaPi * (iota * iota) - kAlphaPi * iota * (ci + cq) / 2) / ((rhoi + rq) / 2);
    }
    const int set = crit <= 0. ? 1 : 0;
    d.piv[g] = set ? pi : 0.0;
    d.pflg[g] = set;
    d.pidx[j * N + p] = g;
    const int nq = d.nnb[q];
    if (nq > kNbrMax) {
      atomicOr(&d.dyn[k].err, kErrCapacity);
      continue;
    }
    for (int jj = 0; jj < nq; ++jj) {
      if (d.nbr[jj * N + q] == p) {
        d.pidx[jj * N + q] = g;
        break;
      }
    }
  }
}

// total pair ranks of this evaluation (the scan's last entry)
__device__ __forceinline__ int pair_total(const DevState &d) { return d.pair_start[d.ctot]; }

// per block of kPairBlk ranks: the last rank that sets pi_ij (-1: none)
__global__ void __launch_bounds__(kPairBlk) k_avlast(DevState d) {
  const int g = blockIdx.x * kPairBlk + threadIdx.x;
  if ((int)blockIdx.x * kPairBlk >= d.pcap) return;
  const int tot = min(pair_total(d), d.pcap);
  const int v = (g < tot && d.pflg[g]) ? g : -1;
  __shared__ int wm[kPairBlk / 64];
  int w = v;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) w = max(w, __shfl_xor(w, o, 64));
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    int r = wm[0];
    for (int i = 1; i < kPairBlk / 64; ++i) r = max(r, wm[i]);
    d.pblast[blockIdx.x] = r;
  }
}

// the instance whose pair ranks hold g: the last with pair_start[c_off] <= g
__device__ __forceinline__ int pair_instance(const DevState &d, int g) {
  int a = 0, b = d.ninst - 1;
  while (a < b) {
    const int mid = (a + b + 1) >> 1;
    if (d.pair_start[d.meta[mid].c_off] <= g) a = mid; else b = mid - 1;
  }
  return a;
}

__global__ void __launch_bounds__(kPairBlk) k_avres(DevState d) {
  const int b0 = blockIdx.x * kPairBlk;
  if (b0 >= d.pcap) return;
  const int g = b0 + threadIdx.x;
  const int tot = min(pair_total(d), d.pcap);
  const bool in = g < tot;
  const int k = pair_instance(d, in ? g : max(min(tot, d.pcap) - 1, 0));
  const int g0 = d.pair_start[d.meta[k].c_off];
  const int fl = in ? d.pflg[g] : 0;
  // inclusive max-scan of the setting ranks over the block
  __shared__ int wm[kPairBlk / 64];
  __shared__ int prev;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int v = fl ? g : -1;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(v, o, 64);
    if (lane >= o) v = max(v, t);
  }
  if (lane == 63) wm[wid] = v;
  if (threadIdx.x == 0) {
    // the last setting rank before the block, down to the start of the
    // instance of its first rank
    const int kf = pair_instance(d, min(b0, max(tot - 1, 0)));
    const int gf = d.pair_start[d.meta[kf].c_off];
    int r = -1;
    for (int bb = (int)blockIdx.x - 1; bb >= 0 && (bb + 1) * kPairBlk > gf; --bb) {
      r = d.pblast[bb];
      if (r >= 0) break;
    }
    prev = r;
  }
  __syncthreads();
  for (int i = 0; i < wid; ++i) v = max(v, wm[i]);
  if (!in || fl || d.meta[k].avis == 0) return;
  int src = max(v, prev);
  d.piv[g] = src >= g0 ? d.piv[src] : 0.0;
}

// The end-of-step pair pass of instances with Benz's smoothing length
// (smoothing_length.f90:75-109) or the Monaghan time step's viscous CFL
// terms (time_step.f90:89-107): both sum over the last evaluation's pairs,
// in the order the reference's list reaches the particle, with that
// evaluation's gradients and the step's final velocities and densities.
// Benz: drhodt = the persistent sum plus this step's terms (the reference
// never resets it without continuity density; with it, the last density
// call's sum alone), hsml += -hsml drhodt / 2 / rho dt, undone where it
// turns negative.  MODE 0: after the corrector or the leap-frog's second
// evaluation (evaluated state T, final state S; the new hsml goes to S and
// this kernel writes the block's partials); MODE 1: after the Euler
// evaluation, before its in-place update (final velocities and density from
// the tendencies as k_euler forms them; the new hsml and the CFL term go to
// hnew / cfl for k_euler).
template <int MODE>
__global__ void __launch_bounds__(kTPB) k_endpass(DevState d) {
  if (blockIdx.x >= (unsigned)d.nblk) return;
  const BlockEntry be = d.blk[blockIdx.x];
  const int k = be.inst;
  const InstMeta &m = d.meta[k];
  if (!m.endp) return;   // (block-uniform)
  const InstDyn &y = d.dyn[k];
  const bool ok = (int)threadIdx.x < be.nloc;
  const bool live = ok && !inst_off(y);
  const int self = be.pos + (ok ? (int)threadIdx.x : 0);
  const int cs = be.cs;
  const int eb = MODE == 0 ? 1 : 0;   // buffer of the evaluated state
  const double *X = d.x[2 * cs + eb], *Y = d.y[2 * cs + eb], *HS = d.hsml[2 * cs + eb];
  const double *VX = d.vx[2 * cs + 0], *VY = d.vy[2 * cs + 0];
  // masses: Benz's own sum (den_alg != 2) follows mass_evolution and takes
  // the new ones; a continuity rate is the last density call's, which
  // precedes it (time_integration.f90:128-137, 262-271, 372-380): the
  // evaluated state's
  const bool newm = be.tmass && m.den_alg != 2;
  const double *MS = (MODE == 0 && be.tmass && !newm) ? d.masst[cs] : d.mass[cs];
  const size_t N = (size_t)d.ntot;
  const double dt = y.dt;
  double px = 0., py = 0., hsn = 1.0, sp = 0., dh = 0., fo = 0., fc = 0.;
  if (live) {
    // the final velocity of slot t
    auto vel = [&](int t, double &u, double &v) {
      u = VX[t];
      v = VY[t];
      if (MODE == 1) {
        u = u + d.tend[2 * N + t] * dt;
        v = v + d.tend[3 * N + t] * dt;
      }
    };
    // the mass of slot t the sums take (Euler: the new masses from the
    // tendencies, as k_euler forms them)
    auto mass_of = [&](int t) {
      if (MODE == 1 && newm) {
        double An = d.A[2 * cs][t] + d.tend[5 * N + t] * dt;
        if (An > 1) An = 1;
        return area_mass(An, d.h[2 * cs][t] + d.tend[4 * N + t] * dt, HS[t]);
      }
      return MS[t];
    };
    const int skf = m.skf;
    const double ksc = kern_scale(skf);
    const double xi = X[self], yi = Y[self], hi = HS[self];
    double ui, vi;
    vel(self, ui, vi);
    // the final density
    double rn;
    if (MODE == 0) {
      rn = d.rho[2 * cs + 0][self];
    } else {
      rn = m.den_alg != 3 ? d.dsum[self]
                          : (d.h[2 * cs][self] + d.tend[4 * N + self] * dt) * kRhoIce;
    }
    const bool benz = m.sle == 1;
    const bool keep = benz && m.den_alg != 2;   // the persistent drhodt
    const int ri = m.p_off + d.orig[cs][self];
    double drho = keep ? d.drp[ri] : 0.0;
    double divv = 0.;
    auto pair = [&](int t, double dx, double dy, double r, double mh) {
      double uq, vq;
      vel(t, uq, vq);
      const double gg = kern_grad(skf, r, mh);
      const double dw0 = gg * (dx / r) * ksc, dw1 = gg * (dy / r) * ksc;
      const double w1 = (ui - uq) * dw0, w4 = (vi - vq) * dw1;
      const double mq = mass_of(t);
      drho = drho + mq * (w1 + w4);
      divv = divv - mq * (w1 + w4) / rn;
    };
    const int nn = d.nnb[self];
    if (nn <= kNbrMax) {
      for (int j = 0; j < nn; ++j) {
        const int t = d.nbr[j * N + self];
        const double dx = xi - X[t], dy = yi - Y[t];
        pair(t, dx, dy, sqrt(dx * dx + dy * dy), (hi + HS[t]) / 2.);
      }
    } else {
      const int W = y.W;
      const int c = d.cellid[cs][self];
      const int cy = c / W, cx = c - cy * W;
      const Ranges R = neighbour_ranges(d.cell_start + m.c_off, W, cx, cy);
      for (int g = 0; g < 6; ++g) {
        for (int t = R.a[g]; t < R.b[g]; ++t) {
          if (t == self) continue;
          const double dx = xi - X[t], dy = yi - Y[t];
          const double mh = (hi + HS[t]) / 2.;
          double r;
          if (!in_support(dx * dx + dy * dy, mh, r)) continue;
          pair(t, dx, dy, r, mh);
        }
      }
    }
    if (benz) {
      const double dhdt = -(hi * drho / 2 / rn);
      hsn = hi + dhdt * dt;
      if (hsn < 0) hsn = hsn - dhdt * dt;
      if (keep) d.drp[ri] = drho;
    } else if (m.sle == 2) {
      if (MODE == 0) {
        hsn = d.hsml[2 * cs + 0][self];          // pass B's Marquis_hsml
      } else {
        hsn = sqrt(mass_of(self) / rn) * kHsmlMult;   // (smoothing_length.f90:136-142)
        if (hsn > m.max_hsml) hsn = m.max_hsml;
      }
    } else {
      hsn = hi;
    }
    if (m.tse == 1) {
      const double ax = d.tend[2 * N + self], ay = d.tend[3 * N + self];
      monaghan_terms(hsn, rn, ax, ay, fo, fc);
      if (m.avdt) fc = monaghan_cfl_av(hsn, rn, divv);
    }
    if (MODE == 0) {
      if (benz) d.hsml[2 * cs + 0][self] = hsn;
      px = d.x[2 * cs + 0][self];
      py = d.y[2 * cs + 0][self];
      const double ue = d.vx[2 * cs + 1][self], ve = d.vy[2 * cs + 1][self];
      sp = sqrt(ue * ue + ve * ve);
      dh = fabs(hsn - hi);
    } else {
      d.hnew[self] = hsn;
      d.cfl[self] = fc;
    }
  }
  if (MODE == 0) block_partials<kParts>(d, k, live, px, py, hsn, sp, dh, fo, fc);
}

// ------------------------------------------------------- pair-list export

// Particle pairs owned by p at their ranks in the reference's order
// (nnps.f90:259-411): p_i, p_j 1-based, r_ij, hsml_ij, and optionally the
// kernel values w and gradients dwdx (2 per pair, dx = x(p_i) - x(p_j)).
__global__ void __launch_bounds__(kTPB) k_export_pairs(DevState d, int buf, int cap, int *pi,
                                                       int *pj, double *rij, double *hij,
                                                       double *w, double *dwdx) {
  int p = 0, k = 0;
  if (!particle_of(d, p, k)) return;
  const InstDyn &y = d.dyn[k];
  const InstMeta &m = d.meta[k];
  if (inst_off(y)) return;
  const int cs = cur_set(y);
  const int s = p;
  const int oc = d.own_cnt[s];
  if ((oc & 0xfffff) == 0) return;
  const int c = d.cellid[cs][p];
  const int pbase = d.pair_start[m.c_off];
  const int base = d.pair_start[m.c_off + c] - pbase;
  int rI = base + (d.moff[s] & 0xffff);
  int rX = base + d.cell_intra[m.c_off + c] + (d.moff[s] >> 16);
  const int W = y.W, cy = c / W, cx = c - cy * W;
  const double xi = d.x[2 * cs + buf][s], yi = d.y[2 * cs + buf][s], hi = d.hsml[2 * cs + buf][s];
  const Ranges R = neighbour_ranges(d.cell_start + m.c_off, W, cx, cy);
  for (int g = 2; g < 6; ++g) {
    for (int t = R.a[g]; t < R.b[g]; ++t) {
      if (g == 2 && t <= s) continue;
      const double dx = xi - d.x[2 * cs + buf][t], dy = yi - d.y[2 * cs + buf][t];
      const double mh = (hi + d.hsml[2 * cs + buf][t]) / 2.;
      double r;
      if (!in_support(dx * dx + dy * dy, mh, r)) continue;
      const int q = t;
      int &rank = (g == 2) ? rI : rX;
      if (rank < cap) {
        pi[rank] = d.orig[cs][p] + 1;
        pj[rank] = d.orig[cs][q] + 1;
        rij[rank] = r;
        hij[rank] = mh;
        if (w) w[rank] = kern_w(m.skf, r, mh);
        if (dwdx) {
          const double gg = kern_grad(m.skf, r, mh), ksc = kern_scale(m.skf);
          dwdx[2 * rank] = gg * (dx / r) * ksc;
          dwdx[2 * rank + 1] = gg * (dy / r) * ksc;
        }
      }
      ++rank;
    }
  }
}

// Boundary pairs (nnps.f90:453-492) at their ranks: bp_i (boundary), bp_j
// (particle), br_ij.  One thread per boundary particle of instance 0.
__global__ void __launch_bounds__(kTPB) k_export_bpairs(DevState d, int buf, int bcap,
                                                        int *bpi, int *bpj, double *brij) {
  const int k = 0;
  const InstMeta &m = d.meta[k];
  const InstDyn &y = d.dyn[k];
  const int ss = cur_set(y);
  int b = blockIdx.x * kTPB + threadIdx.x;
  if (b >= m.nb || inst_off(y)) return;
  const int bg = m.b_off + b;
  const double bin = y.bin;
  const int W = y.W;
  const double bx = d.xb[bg], by = d.yb[bg];
  const int cxb = nint(bx / bin) - y.xlo + 1, cyb = nint(by / bin) - y.ylo + 1;
  const int *cs = d.cell_start + m.c_off;
  int rank = d.bbase[bg];
  for (int s9 = 0; s9 < 9; ++s9) {
    int c = (cyb + s9 / 3 - 1) * W + (cxb + s9 % 3 - 1);
    for (int s = cs[c]; s < cs[c + 1]; ++s) {
      const int q = s;
      double dx = bx - d.x[2 * ss + buf][s], dy = by - d.y[2 * ss + buf][s];
      double dd = sqrt(dx * dx + dy * dy);
      if (dd <= kBLength) {
        if (rank < bcap) {
          bpi[rank] = b + 1;
          bpj[rank] = d.orig[ss][q] + 1;
          brij[rank] = dd;
        }
        ++rank;
      }
    }
  }
}

// ------------------------------------------ one instance in caller memory

// The block table of one instance of n particles at packed offset 0, built
// on the device (the table the host builds for a packed layout): physical
// block j takes logical block x q + min(x, r) + j / 8, x = j mod 8, q and
// r the quotient and remainder of nblk / 8, so that the workgroups each XCD
// receives round-robin cover a contiguous run of slots.
__global__ void __launch_bounds__(kTPB) k_single_blocks(BlockEntry *blk, int n, int nblk) {
  const int j = blockIdx.x * kTPB + threadIdx.x;
  if (j >= nblk) return;
  const int q = nblk / 8, r = nblk % 8, x = j % 8;
  const int lb = x * q + min(x, r) + j / 8;
  BlockEntry e;
  e.inst = 0;
  e.first = lb * kTPB;
  e.lid = lb;
  e.nloc = min(n - e.first, kTPB);
  e.pos = e.first;
  e.cs = 0;
  e.tmass = 0;
  e.pad = 0;
  blk[j] = e;
}

// A call's state in the caller's layout (positions, velocities and boundary
// positions as (x, y) pairs per particle, copied into device memory as they
// are) into state set 0, slot = particle index.
__global__ void __launch_bounds__(kTPB) k_single_unpack(DevState d, const double2 *x2,
                                                        const double2 *v2, const double2 *xb2,
                                                        int n, int nb) {
  const int i = blockIdx.x * kTPB + threadIdx.x;
  if (i < n) {
    const double2 p = x2[i], v = v2[i];
    d.x[0][i] = p.x;
    d.y[0][i] = p.y;
    d.vx[0][i] = v.x;
    d.vy[0][i] = v.y;
    d.orig[0][i] = i;
  }
  if (i < nb) {
    const double2 b = xb2[i];
    d.xb[i] = b.x;
    d.yb[i] = b.y;
  }
}

// The results of a single evaluation from slot order into the caller's
// particle order and layout, at fixed offsets of out (in units of n
// doubles): dxdt (2, n) at 0, dvxdt (2, n) at 2, dhdt at 4, dAdt at 5,
// eps3 (3, n) at 6, zeta at 9, eta at 10, p at 11, sig3 (3, n) at 12.
__global__ void __launch_bounds__(kTPB) k_single_pack(DevState d, double *out, int n,
                                                      int want_t, int want_d) {
  const int s = blockIdx.x * kTPB + threadIdx.x;
  if (s >= n) return;
  const InstDyn &y = d.dyn[0];
  const int r = d.orig[y.sel ^ y.rb][s];
  const size_t N = (size_t)d.ntot, m = (size_t)n;
  if (want_t) {
    const double *T = d.tend;
    out[2 * r] = T[s];
    out[2 * r + 1] = T[N + s];
    out[2 * m + 2 * r] = T[2 * N + s];
    out[2 * m + 2 * r + 1] = T[3 * N + s];
    out[4 * m + r] = T[4 * N + s];
    out[5 * m + r] = T[5 * N + s];
  }
  if (want_d) {
    const double *D = d.diag, *S = d.sig_out;
    for (int c = 0; c < 3; ++c) {
      out[6 * m + 3 * r + c] = D[c * N + s];
      out[12 * m + 3 * r + c] = S[c * N + s];
    }
    out[9 * m + r] = D[3 * N + s];
    out[10 * m + r] = D[4 * N + s];
    out[11 * m + r] = D[5 * N + s];
  }
}

// --------------------------------------------------------------- launches

static inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }
// grid rounded up to a multiple of the 8 XCDs (see particle_of)
// Every launch goes through launch_k: while g_kernel_events is set (the
// engine's profiler), the launch records its kernel's own start and end
// (hipExtLaunchKernelGGL: timestamps of the dispatch itself, the figures a
// kernel trace reports), one event pair per kernel.
thread_local KernelEvents *g_kernel_events = nullptr;

template <class K, class... A>
static void launch_k(K kern, long grid, int block, hipStream_t st, A... args) {
  KernelEvents *ke = g_kernel_events;
  if (ke) {
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess) {
      hipExtLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(block), 0, st, a, b, 0, args...);
      ke->ev.emplace_back(a, b);
      return;
    }
    if (a) (void)hipEventDestroy(a);
  }
  kern<<<(unsigned)grid, block, 0, st>>>(args...);
}

static inline int g8(long blocks) { return (int)((std::max<long>(blocks, 1) + 7) / 8 * 8); }

void launch_partial(const DevState &d, int buf, hipStream_t st) {
  launch_k(k_partial, g8(d.nblk), kTPB, st, d, buf);
}
void launch_finalize(const DevState &d, int end_of_step, hipStream_t st) {
  launch_k(k_finalize, g8(d.ninst), kTPB, st, d, end_of_step);
}
void launch_stats_reset(const DevState &d, hipStream_t st) {
  launch_k(k_stats_reset, cdiv(std::max(d.ninst, 1), kTPB), kTPB, st, d);
}
void launch_fold(const DevState &d, hipStream_t st) {
  launch_k(k_fold, cdiv(std::max(d.ninst, 1), kTPB), kTPB, st, d);
}
void launch_count(const DevState &d, int buf, hipStream_t st) {
  launch_k(k_count, g8(d.nblk), kTPB, st, d, buf);
}
void launch_rebuild(const DevState &d, int buf, int *tmp, hipStream_t st) {
  const int nb = std::max(cdiv(d.ctot, kScanItems), 1);
  // one block per block-table entry (or scan block) up to four per CU, so
  // that every block of the grid is resident at once
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus < 1)
    cus = kRebuildBlocks / 4;
  const long grid = std::min<long>({(long)g8(std::max(d.nblk, nb)), (long)kRebuildBlocks,
                                    4L * cus});
  launch_k(k_rebuild, grid, kTPB, st, d, tmp, nb, buf);
}
void launch_scan(int *in, int L, int *tmp, int *out, int zero_in, const int *gate,
                 hipStream_t st) {
  int nb = cdiv(L, kScanItems);
  if (nb < 1) nb = 1;
  launch_k(k_scan1, g8(nb), kTPB, st, in, L, tmp, nb, gate);
  launch_k(k_scan2, g8(nb), kTPB, st, in, L, tmp, out, zero_in, nb, gate);
}
// The tuned variants exist for the reference's default switch values; the
// variants for other values run the default batch sizes.
template <bool MIXED>
static void launch_passA_t(const DevState &d, int buf, int want_diag, int ba, hipStream_t st) {
  if (ba == 1)
    launch_k(k_passA<MIXED, false, 1>, g8(d.nblk), kTPB, st, d, buf, want_diag);
  else if (ba == 3)
    launch_k(k_passA<MIXED, false, 3>, g8(d.nblk), kTPB, st, d, buf, want_diag);
  else if (ba == 4)
    launch_k(k_passA<MIXED, false, 4>, g8(d.nblk), kTPB, st, d, buf, want_diag);
  else
    launch_k(k_passA<MIXED, false, 2>, g8(d.nblk), kTPB, st, d, buf, want_diag);
}
void launch_passA(const DevState &d, int buf, int mixed, int xsw, int want_diag, int ba,
                  hipStream_t st) {
  if (mixed && xsw)
    launch_k(k_passA<true, true, kBatchA>, g8(d.nblk), kTPB, st, d, buf, want_diag);
  else if (mixed)
    launch_passA_t<true>(d, buf, want_diag, ba, st);
  else if (xsw)
    launch_k(k_passA<false, true, kBatchA>, g8(d.nblk), kTPB, st, d, buf, want_diag);
  else
    launch_passA_t<false>(d, buf, want_diag, ba, st);
}
void launch_pairtot(const DevState &d, hipStream_t st) {
  launch_k(k_pairtot, g8(cdiv(d.ctot, kTPB)), kTPB, st, d);
}
void launch_ranks(const DevState &d, int buf, hipStream_t st) {
  launch_k(k_ranks, g8(d.ninst), kRankThreads, st, d, buf);
}
template <bool MIXED, bool XSW, int phase>
static void launch_passB_b(const DevState &d, int bb, hipStream_t st) {
  constexpr int def = MIXED ? kBatchBMixed : kBatchBFP64;
  // tuned variants: the predictor and corrector of the default switches
  if (!XSW && phase <= 1 && bb == 2)
    launch_k(k_passB<MIXED, XSW, phase, (!XSW && phase <= 1) ? 2 : def>, g8(d.nblk), kTPB, st, d);
  else if (!XSW && phase <= 1 && bb == 3)
    launch_k(k_passB<MIXED, XSW, phase, (!XSW && phase <= 1) ? 3 : def>, g8(d.nblk), kTPB, st, d);
  else if (!XSW && phase <= 1 && bb == 4)
    launch_k(k_passB<MIXED, XSW, phase, (!XSW && phase <= 1) ? 4 : def>, g8(d.nblk), kTPB, st, d);
  else
    launch_k(k_passB<MIXED, XSW, phase, def>, g8(d.nblk), kTPB, st, d);
}
template <bool MIXED, bool XSW>
static void launch_passB_t(const DevState &d, int phase, int bb, hipStream_t st) {
  if (phase == 0)
    launch_passB_b<MIXED, XSW, 0>(d, bb, st);
  else if (phase == 1)
    launch_passB_b<MIXED, XSW, 1>(d, bb, st);
  else if (phase == 2)
    launch_passB_b<MIXED, XSW, 2>(d, bb, st);
  else if (phase == 3)
    launch_passB_b<MIXED, XSW, 3>(d, bb, st);
  else
    launch_passB_b<MIXED, XSW, 4>(d, bb, st);
}
void launch_euler(const DevState &d, hipStream_t st) {
  launch_k(k_euler, g8(d.nblk), kTPB, st, d);
}
void launch_condens(const DevState &d, int mode, hipStream_t st) {
  if (mode == 0)
    launch_k(k_condens<0>, g8(d.nblk), kTPB, st, d);
  else if (mode == 1)
    launch_k(k_condens<1>, g8(d.nblk), kTPB, st, d);
  else
    launch_k(k_condens<2>, g8(d.nblk), kTPB, st, d);
}
void launch_av(const DevState &d, int buf, hipStream_t st) {
  launch_k(k_avpair, g8(d.nblk), kTPB, st, d, buf);
  const int nbp = g8(cdiv(std::max(d.pcap, 1), kPairBlk));
  launch_k(k_avlast, nbp, kPairBlk, st, d);
  launch_k(k_avres, nbp, kPairBlk, st, d);
}
void launch_endpass(const DevState &d, int mode, hipStream_t st) {
  if (mode == 0)
    launch_k(k_endpass<0>, g8(d.nblk), kTPB, st, d);
  else
    launch_k(k_endpass<1>, g8(d.nblk), kTPB, st, d);
}
void launch_passB(const DevState &d, int mixed, int xsw, int phase, int bb, hipStream_t st) {
  if (mixed && xsw)
    launch_passB_t<true, true>(d, phase, bb, st);
  else if (mixed)
    launch_passB_t<true, false>(d, phase, bb, st);
  else if (xsw)
    launch_passB_t<false, true>(d, phase, bb, st);
  else
    launch_passB_t<false, false>(d, phase, bb, st);
}
void launch_export_pairs(const DevState &d, int buf, int cap, int *pi, int *pj, double *rij,
                         double *hij, double *w, double *dwdx, hipStream_t st) {
  launch_k(k_export_pairs, g8(d.nblk), kTPB, st, d, buf, cap, pi, pj, rij, hij, w, dwdx);
}
void launch_export_bpairs(const DevState &d, int buf, int nb, int bcap, int *bpi, int *bpj,
                          double *brij, hipStream_t st) {
  if (nb > 0) launch_k(k_export_bpairs, g8(cdiv(nb, kTPB)), kTPB, st, d, buf, bcap, bpi, bpj, brij);
}
void launch_single_blocks(const DevState &d, int n, hipStream_t st) {
  launch_k(k_single_blocks, cdiv(std::max(d.nblk, 1), kTPB), kTPB, st, d.blk, n, d.nblk);
}
void launch_single_unpack(const DevState &d, const double *x2, const double *v2,
                          const double *xb2, int n, int nb, hipStream_t st) {
  launch_k(k_single_unpack, cdiv(std::max(std::max(n, nb), 1), kTPB), kTPB, st, 
      d, (const double2 *)x2, (const double2 *)v2, (const double2 *)xb2, n, nb);
}
void launch_single_pack(const DevState &d, double *out, int n, int want_t, int want_d,
                        hipStream_t st) {
  launch_k(k_single_pack, cdiv(std::max(n, 1), kTPB), kTPB, st, d, out, n, want_t, want_d);
}

}  // namespace simp
