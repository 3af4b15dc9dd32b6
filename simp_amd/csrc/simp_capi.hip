// simp_capi.hip -- the extern "C" boundary declared in include/simp.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/simp.h"
#include "simp_engine.h"

using simp::Engine;
using simp::InstanceInput;

struct simp_handle {
  simp_config cfg;
  Engine *eng = nullptr;
  int dev = 0;                // the HIP device current at simp_create
  bool uploaded = false;
  int n = 0, nb = 0;
  int g_nx = 0, g_ny = 0;
  double g_x0 = 0, g_y0 = 0, g_dx = 1, g_dy = 1;
  std::vector<double> g_u, g_v;
  long grid_ver = 0;          // simp_set_wind_grid calls so far
  long grid_ver_loaded = -1;  // the version the resident instance was uploaded with
  double step_dt = 0.0;       // simp_set_module_state: dt, boundary normals
  std::vector<double> bnormal;
};

// ---- device memory of simp_single_step_noalloc ----
// One array per GPU the library runs on, held until the process ends: the
// hsml_ij store and its pair slots, kernel_zero, then the layout of the
// call's instance and the staging of its inputs and results
// (Engine::single_in_memory).  In host code `scratch` names a host-side
// stand-in, not device memory: only the address hipGetSymbolAddress
// returns for the current device is ever passed on.
constexpr size_t kScratchBytes = (size_t)SIMP_NOALLOC_SCRATCH_BYTES;
constexpr long kScratchStore = 1L << 18;   // hsml_ij store entries
constexpr long kScratchKz = 1L << 19;      // kernel_zero entries (particles)
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
struct ScratchDevice {
  bool looked_up = false;
  void *ptr = nullptr;                // device address of `scratch` (null: lookup failed)
  const simp_handle *last = nullptr;  // the handle whose call laid out its content last
  long kz_valid = 0;                  // particles of that call (kernel_zero entries kept)
};
ScratchDevice g_scratch[kMaxDevices];
std::mutex g_scratch_mu;

// Once per device, with that device current (at simp_create).
void scratch_lookup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  ScratchDevice &sd = g_scratch[dev];
  if (sd.looked_up) return;
  sd.looked_up = true;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) sd.ptr = p;
  else (void)hipGetLastError();   // the no-allocation entry point reports it
}

// A handle whose device state lies in its GPU's array can read it back only
// while no other handle's call has laid out its own there.
bool scratch_state_lost(const simp_handle *h) {
  return h->eng->in_caller_memory() &&
         (h->dev < 0 || h->dev >= kMaxDevices || g_scratch[h->dev].last != h);
}
}  // namespace

static int map_rc(int rc) {
  switch (rc) {
    case 0: return SIMP_OK;
    case -1: return SIMP_E_ARG;
    case -3: return SIMP_E_CAPACITY;
    case -4: return SIMP_E_NAN;
    case -5: return SIMP_E_STATE;
    case -6: return SIMP_E_NOMEM;
    default: return SIMP_E_HIP;
  }
}

extern "C" {

void simp_config_default(simp_config *c) {
  if (!c) return;
  *c = simp_config{};
  c->version = SIMP_CONFIG_VERSION;
  c->precision = SIMP_FP64;
  c->int_algorithm = 2;
  c->nnps_algorithm = 2;
  c->sle = 2;
  c->tse = 2;
  c->rheology = 1;
  c->bfa = 1;
  c->eos_alg = 3;
  c->pa_sph = 1;
  c->skf = 6;
  c->den_alg = 3;
  c->wind_mode = SIMP_WIND_UNIFORM;
  c->wind[0] = 0.0;
  c->wind[1] = (double)(-5.3f);
  c->current[0] = c->current[1] = 0.0;
  c->max_hsml = 0.0;
  c->artificial_visc = 0;
  c->art_visc_alg = 1;
  c->xsph_regularization = 0;
  c->boundary_friction = 0;
  c->boundary_pressure = 1;
  c->pressure_force = 1;
  c->wind_force = 1;
  c->water_force = 1;
  c->viscosity = 1;
}

const char *simp_strerror(int s) {
  switch (s) {
    case SIMP_OK: return "ok";
    case SIMP_E_ARG: return "invalid argument";
    case SIMP_E_HIP: return "HIP runtime failure";
    case SIMP_E_CAPACITY: return "capacity exceeded";
    case SIMP_E_NAN: return "non-finite position or smoothing length";
    case SIMP_E_UNSUPPORTED: return "switch value not implemented";
    case SIMP_E_NOMEM: return "device allocation failed";
    case SIMP_E_STATE: return "call out of order";
    default: return "unknown status";
  }
}

static int validate(const simp_config *c) {
  if (!c || c->version != SIMP_CONFIG_VERSION) return SIMP_E_ARG;
  if (c->precision != SIMP_FP64 && c->precision != SIMP_MIXED) return SIMP_E_UNSUPPORTED;
  if (c->int_algorithm < 1 || c->int_algorithm > 3) return SIMP_E_UNSUPPORTED;
  if (c->nnps_algorithm != 2 || c->sle < 0 || c->sle > 2) return SIMP_E_UNSUPPORTED;
  if (c->tse < 0 || c->tse > 2) return SIMP_E_UNSUPPORTED;
  if ((c->rheology != 1 && c->rheology != 2) || c->bfa < 1 || c->bfa > 3)
    return SIMP_E_UNSUPPORTED;
  if (c->art_visc_alg != 1 && c->art_visc_alg != 2) return SIMP_E_UNSUPPORTED;
  if (c->eos_alg < 1 || c->eos_alg > 3) return SIMP_E_UNSUPPORTED;
  if (c->pa_sph != 1 && c->pa_sph != 2) return SIMP_E_UNSUPPORTED;
  if (c->skf < 1 || c->skf > 8 || c->den_alg < 1 || c->den_alg > 3) return SIMP_E_UNSUPPORTED;
  if (c->wind_mode != SIMP_WIND_UNIFORM && c->wind_mode != SIMP_WIND_GRID) return SIMP_E_UNSUPPORTED;
  return SIMP_OK;
}

int simp_create(const simp_config *cfg, simp_handle **out) {
  if (!out) return SIMP_E_ARG;
  *out = nullptr;
  int v = validate(cfg);
  if (v) return v;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SIMP_E_HIP;
  simp_handle *h = new (std::nothrow) simp_handle;
  if (!h) return SIMP_E_NOMEM;
  h->cfg = *cfg;
  if (hipGetDevice(&h->dev) != hipSuccess) {
    delete h;
    return SIMP_E_HIP;
  }
  scratch_lookup(h->dev);
  h->eng = new (std::nothrow) Engine(cfg->precision);
  if (!h->eng || !h->eng->stream()) {
    delete h->eng;
    delete h;
    return SIMP_E_HIP;
  }
  h->eng->set_scheme(cfg->int_algorithm);
  *out = h;
  return SIMP_OK;
}

int simp_destroy(simp_handle *h) {
  if (!h) return SIMP_E_ARG;
  if (h->dev >= 0 && h->dev < kMaxDevices && g_scratch[h->dev].last == h)
    g_scratch[h->dev].last = nullptr;
  delete h->eng;
  delete h;
  return SIMP_OK;
}

int simp_set_wind_grid(simp_handle *h, int nx, int ny, double x0, double y0, double dx,
                       double dy, const double *u, const double *v) {
  if (!h || nx < 2 || ny < 2 || !u || !v || !(dx > 0) || !(dy > 0)) return SIMP_E_ARG;
  h->g_nx = nx; h->g_ny = ny;
  h->g_x0 = x0; h->g_y0 = y0; h->g_dx = dx; h->g_dy = dy;
  h->g_u.assign(u, u + (size_t)nx * ny);
  h->g_v.assign(v, v + (size_t)nx * ny);
  h->grid_ver += 1;
  return SIMP_OK;
}

// Without a boundary force the reference finds no boundary pairs
// (nnps.f90:440) and never reads the boundary particles.
static void drop_unused_boundary(InstanceInput &in) {
  if (in.boundary_pressure || in.boundary_friction) return;
  in.nb = 0;
  in.xb = in.massb = in.bnormal = nullptr;
}

static int fill_input(simp_handle *h, InstanceInput &in) {
  const simp_config &c = h->cfg;
  in.tse = c.tse;
  in.wind_mode = c.wind_mode;
  in.wind[0] = c.wind[0]; in.wind[1] = c.wind[1];
  in.current[0] = c.current[0]; in.current[1] = c.current[1];
  in.max_hsml = c.max_hsml;
  in.den_alg = c.den_alg;
  in.pa_sph = c.pa_sph;
  in.eos_alg = c.eos_alg;
  in.rheology = c.rheology;
  in.skf = c.skf;
  in.sle = c.sle;
  in.xsph = c.xsph_regularization ? 1 : 0;
  in.art_visc = c.artificial_visc ? 1 : 0;
  in.art_visc_alg = c.art_visc_alg;
  in.bfa = c.bfa;
  if (!h->bnormal.empty()) {
    if ((int)(h->bnormal.size() / 2) != in.nb) return SIMP_E_ARG;
    in.bnormal = h->bnormal.data();
  }
  in.boundary_friction = c.boundary_friction ? 1 : 0;
  in.boundary_pressure = c.boundary_pressure ? 1 : 0;
  in.pressure_force = c.pressure_force ? 1 : 0;
  in.wind_force = c.wind_force ? 1 : 0;
  in.water_force = c.water_force ? 1 : 0;
  in.viscosity = c.viscosity ? 1 : 0;
  drop_unused_boundary(in);
  if (c.wind_mode == SIMP_WIND_GRID) {
    if (h->g_nx < 2) return SIMP_E_STATE;
    in.g_nx = h->g_nx; in.g_ny = h->g_ny;
    in.g_x0 = h->g_x0; in.g_y0 = h->g_y0; in.g_dx = h->g_dx; in.g_dy = h->g_dy;
    in.g_u = h->g_u.data(); in.g_v = h->g_v.data();
  }
  if (!(in.max_hsml > 0)) {
    // initialize.f90:191 convention when the caller gives no cap
    double mx = 0;
    for (int i = 0; i < in.n; ++i) mx = std::max(mx, in.hsml[i]);
    in.max_hsml = mx * 10.;
  }
  return SIMP_OK;
}

int simp_single_step(simp_handle *h, int n, const double *x, const double *vx,
                     const double *mass, const double *rho, const double *hsml,
                     const double *hth, const double *A, int nb, const double *xb,
                     const double *massb, double current_time, int itimestep, double *dxdt,
                     double *dvxdt, double *dhdt, double *dAdt, double *eps3, double *zeta,
                     double *eta, double *p, double *sig3, int *nip, int *bnip) {
  if (!h || n < 1 || nb < 0) return SIMP_E_ARG;
  InstanceInput in;
  in.n = n; in.nb = nb;
  in.x = x; in.vx = vx; in.mass = mass; in.rho = rho; in.hsml = hsml; in.hth = hth; in.A = A;
  in.xb = xb; in.massb = massb;
  in.time = current_time;
  in.itimestep = itimestep;
  in.dt = h->step_dt;     // (read by the Wang boundary force only)
  int rc = fill_input(h, in);
  if (rc) return rc;
  // the resident layout of the last call takes the new state when it can
  // (same sizes and switches); otherwise a full upload
  rc = h->eng->reload(in, h->grid_ver == h->grid_ver_loaded);
  if (rc == 1) rc = h->eng->upload({in}, true);
  if (rc) return map_rc(rc);
  h->grid_ver_loaded = h->grid_ver;
  h->uploaded = false;  // the device state now holds this call's inputs
  rc = h->eng->eval_single(eps3 || zeta || eta || p || sig3);
  if (rc) return map_rc(rc);
  rc = h->eng->download_single(dxdt, dvxdt, dhdt, dAdt, eps3, zeta, eta, p, sig3);
  if (rc) return map_rc(rc);
  if (nip || bnip) {
    int a = 0, b = 0;
    rc = h->eng->download_scalars(0, nullptr, nullptr, nullptr, &a, &b, nullptr);
    if (rc) return map_rc(rc);
    if (nip) *nip = a;
    if (bnip) *bnip = nb > 0 ? b : 0;
  }
  return SIMP_OK;
}

int simp_set_module_state(simp_handle *h, double dt, int nb, const double *b_normal) {
  if (!h || nb < 0) return SIMP_E_ARG;
  h->step_dt = dt;
  if (b_normal && nb > 0)
    h->bnormal.assign(b_normal, b_normal + 2 * (size_t)nb);
  else
    h->bnormal.clear();
  return SIMP_OK;
}

int simp_single_step_noalloc(simp_handle *h, int n, const double *x, const double *vx,
                             const double *mass, const double *rho, const double *hsml,
                             const double *hth, const double *A, int nb, const double *xb,
                             const double *massb, double current_time, int itimestep,
                             double *dxdt, double *dvxdt, double *dhdt, double *dAdt,
                             double *eps3, double *zeta, double *eta, double *p, double *sig3,
                             int *nip, int *bnip) {
  if (!h || n < 1 || nb < 0) return SIMP_E_ARG;
  if (h->dev < 0 || h->dev >= kMaxDevices) return SIMP_E_HIP;
  ScratchDevice &sd = g_scratch[h->dev];
  if (!sd.ptr) return SIMP_E_HIP;
  InstanceInput in;
  in.n = n; in.nb = nb;
  in.x = x; in.vx = vx; in.mass = mass; in.rho = rho; in.hsml = hsml; in.hth = hth; in.A = A;
  in.xb = xb; in.massb = massb;
  in.time = current_time;
  in.itimestep = itimestep;
  int rc = fill_input(h, in);
  if (rc) return rc;
  // the regularization, smoothing-length and boundary-force switches need
  // arrays this layout does not carve
  if (in.sle != 2 || in.xsph || in.art_visc || in.bfa != 1 || in.boundary_friction)
    return SIMP_E_UNSUPPORTED;
  // a handle's first call starts the persistent arrays from zero, as a new
  // handle's simp_single_step does; later calls find them as the last call
  // on this GPU left them
  const bool first = !h->eng->in_caller_memory();
  rc = h->eng->single_in_memory((char *)sd.ptr, kScratchBytes, kScratchStore, kScratchKz, in,
                                first, sd.kz_valid);
  if (rc) return map_rc(rc);
  sd.last = h;
  sd.kz_valid = n;
  h->uploaded = false;
  rc = h->eng->eval_single(eps3 || zeta || eta || p || sig3);
  if (rc) return map_rc(rc);
  rc = h->eng->single_in_memory_out(dxdt, dvxdt, dhdt, dAdt, eps3, zeta, eta, p, sig3);
  if (rc) return map_rc(rc);
  if (nip || bnip) {
    int a = 0, b = 0;
    rc = h->eng->download_scalars(0, nullptr, nullptr, nullptr, &a, &b, nullptr);
    if (rc) return map_rc(rc);
    if (nip) *nip = a;
    if (bnip) *bnip = nb > 0 ? b : 0;
  }
  return SIMP_OK;
}

int simp_nnps(simp_handle *h, int n, const double *x, const double *hsml, int nb,
              const double *xb, int cap, int *nip, int *p_i, int *p_j, double *r_ij,
              double *hsml_ij, int bcap, int *bnip, int *bp_i, int *bp_j, double *br_ij) {
  if (!h || n < 1 || nb < 0 || !x || !hsml || !nip || !bnip || cap < 0 || bcap < 0)
    return SIMP_E_ARG;
  if (nb > 0 && !xb) return SIMP_E_ARG;
  std::vector<double> zero2(2 * (size_t)n, 0.0), one(n, 1.0), mb(std::max(nb, 1), 1.0);
  InstanceInput in;
  in.n = n; in.nb = nb;
  in.x = x; in.vx = zero2.data(); in.mass = one.data(); in.rho = one.data();
  in.hsml = hsml; in.hth = one.data(); in.A = one.data();
  in.xb = xb; in.massb = mb.data();
  int rc = fill_input(h, in);
  if (rc) return rc;
  rc = h->eng->upload({in}, true);
  if (rc) return map_rc(rc);
  h->uploaded = false;
  rc = h->eng->export_pairs(cap, nip, p_i, p_j, r_ij, hsml_ij, bcap, bnip, bp_i, bp_j, br_ij);
  return map_rc(rc);
}

int simp_kernel_values(simp_handle *h, int cap, int *nip, int *p_i, int *p_j, double *w,
                       double *kernel_zero) {
  if (!h || cap < 0 || !nip) return SIMP_E_ARG;
  if (h->eng->ninst() != 1 || scratch_state_lost(h)) return SIMP_E_STATE;
  if (kernel_zero && h->cfg.den_alg != 1) return SIMP_E_STATE;
  int bnip = 0;
  int rc = h->eng->export_pairs(cap, nip, p_i, p_j, nullptr, nullptr, 0, &bnip, nullptr,
                                nullptr, nullptr, w);
  // the boundary pairs are not asked for here: only a particle-pair
  // overflow is a capacity error
  if (rc == -3 && *nip <= cap) rc = 0;
  if (rc) return map_rc(rc);
  if (kernel_zero) rc = h->eng->download_kernel_zero(0, kernel_zero);
  return map_rc(rc);
}

int simp_kernel_gradients(simp_handle *h, int cap, int *nip, int *p_i, int *p_j,
                          double *dwdx) {
  if (!h || cap < 0 || !nip) return SIMP_E_ARG;
  if (h->eng->ninst() != 1 || scratch_state_lost(h)) return SIMP_E_STATE;
  int bnip = 0;
  int rc = h->eng->export_pairs(cap, nip, p_i, p_j, nullptr, nullptr, 0, &bnip, nullptr,
                                nullptr, nullptr, nullptr, dwdx);
  // only a particle-pair overflow is a capacity error here
  if (rc == -3 && *nip <= cap) rc = 0;
  return map_rc(rc);
}

static int stats_out(Engine *e, int k, int n, int *maxip, int *minip, double *avgip,
                     int *noip, long long *nip, int *itimestep) {
  long long sum = 0;
  int rc = e->interaction_stats(k, maxip, minip, &sum, noip, nip, itimestep);
  if (rc) return map_rc(rc);
  // real(sumip) / real(ntotal): a REAL(4) division (nnps.f90:651)
  if (avgip) *avgip = (double)((float)sum / (float)n);
  return SIMP_OK;
}

static int tuning_in(const simp_tuning *t, simp::Tuning *out) {
  if (!t || t->passA_batch < 1 || t->passA_batch > 4) return SIMP_E_ARG;
  if (t->passB_batch != 0 && (t->passB_batch < 2 || t->passB_batch > 4)) return SIMP_E_ARG;
  if (!(t->skin > 0.0) || t->skin > 0.2) return SIMP_E_ARG;
  out->passA_batch = t->passA_batch;
  out->passB_batch = t->passB_batch;
  out->skin = (float)t->skin;
  return SIMP_OK;
}

static void tuning_out(const simp::Tuning &t, simp_tuning *out) {
  if (!out) return;
  out->passA_batch = t.passA_batch;
  out->passB_batch = t.passB_batch;
  out->reserved = 0;
  out->skin = (double)t.skin;
}

int simp_set_tuning(simp_handle *h, const simp_tuning *t) {
  if (!h) return SIMP_E_ARG;
  simp::Tuning v;
  int rc = tuning_in(t, &v);
  if (rc) return rc;
  h->eng->set_tuning(v);
  return SIMP_OK;
}

int simp_get_tuning(simp_handle *h, simp_tuning *t) {
  if (!h || !t) return SIMP_E_ARG;
  tuning_out(h->eng->tuning(), t);
  return SIMP_OK;
}

int simp_tune(simp_handle *h, int steps_per_trial, simp_tuning *chosen, double *us_start,
              double *us_best) {
  if (!h || steps_per_trial < 1) return SIMP_E_ARG;
  if (!h->uploaded) return SIMP_E_STATE;
  int rc = h->eng->autotune(steps_per_trial, us_start, us_best);
  if (rc == 0) rc = h->eng->sync_and_check();
  tuning_out(h->eng->tuning(), chosen);
  return map_rc(rc);
}

int simp_set_search(simp_handle *h, int every_evaluation) {
  if (!h || (every_evaluation != 0 && every_evaluation != 1)) return SIMP_E_ARG;
  h->eng->set_force_search(every_evaluation != 0);
  return SIMP_OK;
}

int simp_search_stats(simp_handle *h, long long *evaluations, long long *cell_rebuilds,
                      long long *candidate_rebuilds) {
  if (!h) return SIMP_E_ARG;
  if (h->eng->ninst() != 1 || scratch_state_lost(h)) return SIMP_E_STATE;
  return map_rc(h->eng->search_stats(0, evaluations, cell_rebuilds, candidate_rebuilds));
}

int simp_interaction_stats(simp_handle *h, int *maxip, int *minip, double *avgip, int *noip,
                           long long *nip, int *itimestep) {
  if (!h) return SIMP_E_ARG;
  if (h->eng->ninst() != 1 || scratch_state_lost(h)) return SIMP_E_STATE;
  return stats_out(h->eng, 0, h->eng->meta(0).n, maxip, minip, avgip, noip, nip, itimestep);
}

int simp_upload(simp_handle *h, int n, const double *x, const double *vx, const double *mass,
                const double *rho, const double *hsml, const double *hth, const double *A,
                int nb, const double *xb, const double *massb, double dt,
                double current_time, int itimestep) {
  if (!h || n < 1 || nb < 0) return SIMP_E_ARG;
  InstanceInput in;
  in.n = n; in.nb = nb;
  in.x = x; in.vx = vx; in.mass = mass; in.rho = rho; in.hsml = hsml; in.hth = hth; in.A = A;
  in.xb = xb; in.massb = massb;
  in.dt = dt; in.time = current_time; in.itimestep = itimestep;
  int rc = fill_input(h, in);
  if (rc) return rc;
  rc = h->eng->upload({in}, false);
  if (rc) return map_rc(rc);
  h->grid_ver_loaded = h->grid_ver;
  h->uploaded = true;
  h->n = n;
  h->nb = nb;
  return SIMP_OK;
}

int simp_step_pc(simp_handle *h, int nsteps, int sync) {
  if (!h || nsteps < 0) return SIMP_E_ARG;
  if (!h->uploaded) return SIMP_E_STATE;
  int rc = h->eng->steps(nsteps, true);
  if (rc) return map_rc(rc);
  if (sync) return map_rc(h->eng->sync_and_check());
  return SIMP_OK;
}

int simp_sync(simp_handle *h) {
  if (!h) return SIMP_E_ARG;
  if (!h->uploaded) return SIMP_OK;
  return map_rc(h->eng->sync_and_check());
}

void *simp_stream(simp_handle *h) { return h ? (void *)h->eng->stream() : nullptr; }

// ---- exact snapshots ----
// Blob: "SIMPSNP1", int32 n, nb, itimestep, flags; int64 nq; real8 dt, time;
// real8 x(2n), vx(2n), mass(n), rho(n), hsml(n), h(n), A(n), xb(2nb),
// massb(nb), kernel_zero(n), hsml_ij store(nq); then with flags bit 0 (Benz's
// smoothing length) the persistent drhodt(n), with bit 1 (the Wang boundary
// force) the boundary normals b_normal(2nb); little endian.
static const char kSnapMagic[8] = {'S', 'I', 'M', 'P', 'S', 'N', 'P', '1'};
constexpr int kSnapDrhodt = 1, kSnapNormals = 2;
static long long snap_bytes(long n, long nb, long nq, int flags) {
  return 8 + 16 + 8 + 16 +
         8 * (10 * n + 3 * nb + nq + ((flags & kSnapDrhodt) ? n : 0) +
              ((flags & kSnapNormals) ? 2 * nb : 0));
}
static int snap_flags(const simp_handle *h) {
  return (h->cfg.sle == 1 ? kSnapDrhodt : 0) | (h->cfg.bfa == 2 ? kSnapNormals : 0);
}

long long simp_snapshot_size(simp_handle *h) {
  if (!h) return -SIMP_E_ARG;
  if (!h->uploaded) return -SIMP_E_STATE;
  const long nq = h->eng->store_cap(0);
  if (nq < 0) return -SIMP_E_STATE;
  return snap_bytes(h->n, h->nb, nq, snap_flags(h));
}

int simp_snapshot_save(simp_handle *h, void *buf, long long cap) {
  if (!h || !buf) return SIMP_E_ARG;
  if (!h->uploaded) return SIMP_E_STATE;
  const long n = h->n, nb = h->nb, nq = h->eng->store_cap(0);
  const int flags = snap_flags(h);
  if (nq < 0) return SIMP_E_STATE;
  if (cap < snap_bytes(n, nb, nq, flags)) return SIMP_E_CAPACITY;
  int rc = h->eng->sync_and_check();
  if (rc) return map_rc(rc);
  char *b = (char *)buf;
  std::memcpy(b, kSnapMagic, 8);
  double dt = 0, t = 0;
  int it = 0;
  rc = h->eng->download_scalars(0, &dt, &t, &it, nullptr, nullptr, nullptr);
  if (rc) return map_rc(rc);
  const int32_t hdr[4] = {(int32_t)n, (int32_t)nb, (int32_t)it, (int32_t)flags};
  const int64_t q64 = nq;
  std::memcpy(b + 8, hdr, 16);
  std::memcpy(b + 24, &q64, 8);
  std::memcpy(b + 32, &dt, 8);
  std::memcpy(b + 40, &t, 8);
  double *p = (double *)(b + 48);
  double *x = p, *vx = x + 2 * n, *mass = vx + 2 * n, *rho = mass + n, *hsml = rho + n;
  double *hth = hsml + n, *A = hth + n, *xb = A + n, *massb = xb + 2 * nb, *kz = massb + nb;
  double *hq = kz + n, *drp = hq + nq;
  double *bn = drp + ((flags & kSnapDrhodt) ? n : 0);
  rc = h->eng->download_state(0, x, vx, rho, hsml, hth, A);
  if (!rc)
    rc = h->eng->download_persist(0, mass, xb, massb, kz, hq,
                                  (flags & kSnapDrhodt) ? drp : nullptr,
                                  (flags & kSnapNormals) ? bn : nullptr);
  return map_rc(rc);
}

// Parses and checks a blob's header; the counts are bounded by the doubles
// the blob holds before any size arithmetic, so a forged or truncated
// header cannot overflow it.
static int snap_header(const void *buf, long long len, long *n, long *nb, long *nq, int *it,
                       double *dt, double *t, int *flags) {
  if (!buf || len < 48) return SIMP_E_ARG;
  const char *b = (const char *)buf;
  if (std::memcmp(b, kSnapMagic, 8) != 0) return SIMP_E_ARG;
  int32_t hdr[4];
  int64_t q;
  std::memcpy(hdr, b + 8, 16);
  std::memcpy(&q, b + 24, 8);
  std::memcpy(dt, b + 32, 8);
  std::memcpy(t, b + 40, 8);
  const long long avail = (len - 48) / 8;
  if (hdr[0] < 1 || hdr[1] < 0 || q < 0 || (hdr[3] & ~(kSnapDrhodt | kSnapNormals)))
    return SIMP_E_ARG;
  const long long fixed = 10LL * hdr[0] + 3LL * hdr[1] +
                          ((hdr[3] & kSnapDrhodt) ? (long long)hdr[0] : 0) +
                          ((hdr[3] & kSnapNormals) ? 2LL * hdr[1] : 0);
  if (fixed > avail || q > avail - fixed) return SIMP_E_ARG;
  if (len < snap_bytes(hdr[0], hdr[1], (long)q, hdr[3])) return SIMP_E_ARG;
  *n = hdr[0];
  *nb = hdr[1];
  *nq = (long)q;
  *it = hdr[2];
  *flags = hdr[3];
  return SIMP_OK;
}

int simp_snapshot_info(const void *buf, long long len, int *n, int *nb, int *itimestep,
                       double *current_time) {
  long n_, nb_, nq_;
  int it, fl;
  double dt, t;
  const int rc = snap_header(buf, len, &n_, &nb_, &nq_, &it, &dt, &t, &fl);
  if (rc) return rc;
  if (n) *n = (int)n_;
  if (nb) *nb = (int)nb_;
  if (itimestep) *itimestep = it;
  if (current_time) *current_time = t;
  return SIMP_OK;
}

int simp_snapshot_load(simp_handle *h, const void *buf, long long len) {
  if (!h) return SIMP_E_ARG;
  long n, nb, nq;
  int it, flags;
  double dt, t;
  const int hrc = snap_header(buf, len, &n, &nb, &nq, &it, &dt, &t, &flags);
  if (hrc) return hrc;
  // the blob carries what the handle's switches persist
  if (flags != snap_flags(h)) return SIMP_E_ARG;
  const char *b = (const char *)buf;
  const double *p = (const double *)(b + 48);
  const double *x = p, *vx = x + 2 * n, *mass = vx + 2 * n, *rho = mass + n, *hsml = rho + n;
  const double *hth = hsml + n, *A = hth + n, *xb = A + n, *massb = xb + 2 * nb, *kz = massb + nb;
  const double *hq = kz + n, *drp = hq + nq;
  const double *bn = drp + ((flags & kSnapDrhodt) ? n : 0);
  if (flags & kSnapNormals) h->bnormal.assign(bn, bn + 2 * nb);
  int rc = simp_upload(h, (int)n, x, vx, mass, rho, hsml, hth, A, (int)nb, nb ? xb : nullptr,
                       nb ? massb : nullptr, dt, t, it);
  if (rc) return rc;
  return map_rc(h->eng->upload_persist(0, kz, hq, (long)nq,
                                       (flags & kSnapDrhodt) ? drp : nullptr));
}

int simp_download(simp_handle *h, double *x, double *vx, double *rho, double *hsml, double *hth,
                  double *A, double *dt, double *current_time, int *itimestep) {
  if (!h) return SIMP_E_ARG;
  if (!h->uploaded) return SIMP_E_STATE;
  int rc = h->eng->sync_and_check();
  if (rc) return map_rc(rc);
  rc = h->eng->download_state(0, x, vx, rho, hsml, hth, A);
  if (rc) return map_rc(rc);
  rc = h->eng->download_scalars(0, dt, current_time, itimestep, nullptr, nullptr, nullptr);
  return map_rc(rc);
}

int simp_set_diagnostics(simp_handle *h, int enable) {
  if (!h) return SIMP_E_ARG;
  h->eng->set_diag(enable != 0);
  return SIMP_OK;
}

int simp_download_diagnostics(simp_handle *h, double *eps3, double *zeta, double *eta,
                              double *p, double *sig3) {
  if (!h) return SIMP_E_ARG;
  if (!h->uploaded) return SIMP_E_STATE;
  if (!h->eng->diag() && (eps3 || zeta || eta || p || sig3)) return SIMP_E_STATE;
  return map_rc(h->eng->download_diag(0, eps3, zeta, eta, p, sig3));
}

int simp_profile_steps(simp_handle *h, int nsteps, int maxk, char *buf, int buflen, double *ms,
                       int *counts) {
  if (!h || nsteps < 1 || maxk < 1 || !buf || buflen < 1 || !ms || !counts) return SIMP_E_ARG;
  if (!h->uploaded) return SIMP_E_STATE;
  simp::Profiler prof;
  h->eng->prof = &prof;
  int rc = h->eng->steps(nsteps, false);
  h->eng->prof = nullptr;
  if (rc) return -map_rc(rc);
  if (hipStreamSynchronize(h->eng->stream()) != hipSuccess) return -SIMP_E_HIP;
  std::vector<std::string> names;
  std::map<std::string, std::pair<double, int>> agg;
  // per launch (one or more kernels): the sum of its kernels' durations
  for (auto &r : prof.recs) {
    double t = 0.0;
    for (auto &e : r.ev) {
      float ms = 0.f;
      hipEventElapsedTime(&ms, e.first, e.second);
      t += ms;
      hipEventDestroy(e.first);
      hipEventDestroy(e.second);
    }
    if (!agg.count(r.name)) names.push_back(r.name);
    agg[r.name].first += t;
    agg[r.name].second += 1;
  }
  std::string out;
  int k = 0;
  for (auto &nm : names) {
    if (k >= maxk) break;
    ms[k] = agg[nm].first / agg[nm].second;
    counts[k] = agg[nm].second;
    out += nm;
    out += "\n";
    ++k;
  }
  int len = std::min((int)out.size(), buflen - 1);
  std::copy(out.begin(), out.begin() + len, buf);
  buf[len] = 0;
  return k;
}

/* ------------------------------------------------------------ batches */

struct simp_batch {
  Engine *eng = nullptr;
};

static int fill_from_desc(const simp_instance_desc &d, InstanceInput &in) {
  if (d.n < 1 || d.nb < 0 || !d.x || !d.vx || !d.mass || !d.rho || !d.hsml || !d.hth || !d.A)
    return SIMP_E_ARG;
  if (d.nb > 0 && (!d.xb || !d.massb)) return SIMP_E_ARG;
  if (d.tse < 0 || d.tse > 2) return SIMP_E_UNSUPPORTED;
  if (d.wind_mode != SIMP_WIND_UNIFORM && d.wind_mode != SIMP_WIND_GRID) return SIMP_E_UNSUPPORTED;
  if (d.den_alg < 0 || d.den_alg > 3) return SIMP_E_UNSUPPORTED;
  if (d.pa_sph != 0 && d.pa_sph != 1 && d.pa_sph != 2) return SIMP_E_UNSUPPORTED;
  if (d.eos_alg < 0 || d.eos_alg > 3) return SIMP_E_UNSUPPORTED;
  if (d.rheology < 0 || d.rheology > 2) return SIMP_E_UNSUPPORTED;
  if (d.skf < 0 || d.skf > 8) return SIMP_E_UNSUPPORTED;
  if (d.sle != SIMP_SLE_KEEP && (d.sle < 0 || d.sle > 2)) return SIMP_E_UNSUPPORTED;
  if (d.art_visc_alg < 0 || d.art_visc_alg > 2 || d.bfa < 0 || d.bfa > 3)
    return SIMP_E_UNSUPPORTED;
  if (d.forces_off & ~(SIMP_OFF_BOUNDARY_PRESSURE | SIMP_OFF_PRESSURE_FORCE | SIMP_OFF_WIND |
                       SIMP_OFF_WATER | SIMP_OFF_VISCOSITY))
    return SIMP_E_UNSUPPORTED;
  if (d.wind_mode == SIMP_WIND_GRID &&
      (d.grid_nx < 2 || d.grid_ny < 2 || !d.grid_u || !d.grid_v || !(d.grid_dx > 0) ||
       !(d.grid_dy > 0)))
    return SIMP_E_ARG;
  in.n = d.n; in.nb = d.nb;
  in.x = d.x; in.vx = d.vx; in.mass = d.mass; in.rho = d.rho; in.hsml = d.hsml;
  in.hth = d.hth; in.A = d.A; in.xb = d.xb; in.massb = d.massb;
  in.dt = d.dt; in.time = d.current_time; in.itimestep = d.itimestep;
  in.nsteps = d.nsteps > 0 ? d.nsteps : 0;
  in.tse = d.tse;
  in.den_alg = d.den_alg == 0 ? 3 : d.den_alg;
  in.pa_sph = d.pa_sph == 2 ? 2 : 1;
  in.eos_alg = d.eos_alg == 0 ? 3 : d.eos_alg;
  in.rheology = d.rheology == 2 ? 2 : 1;
  in.skf = d.skf == 0 ? 6 : d.skf;
  in.sle = d.sle == SIMP_SLE_KEEP ? 0 : (d.sle == 1 ? 1 : 2);
  in.xsph = d.xsph ? 1 : 0;
  in.art_visc = d.art_visc ? 1 : 0;
  in.art_visc_alg = d.art_visc_alg == 2 ? 2 : 1;
  in.bfa = d.bfa == 0 ? 1 : d.bfa;
  in.bnormal = d.b_normal;
  in.boundary_friction = d.boundary_friction ? 1 : 0;
  in.boundary_pressure = (d.forces_off & SIMP_OFF_BOUNDARY_PRESSURE) ? 0 : 1;
  in.pressure_force = (d.forces_off & SIMP_OFF_PRESSURE_FORCE) ? 0 : 1;
  in.wind_force = (d.forces_off & SIMP_OFF_WIND) ? 0 : 1;
  in.water_force = (d.forces_off & SIMP_OFF_WATER) ? 0 : 1;
  in.viscosity = (d.forces_off & SIMP_OFF_VISCOSITY) ? 0 : 1;
  drop_unused_boundary(in);
  in.wind_mode = d.wind_mode;
  in.wind[0] = d.wind[0]; in.wind[1] = d.wind[1];
  in.current[0] = d.current[0]; in.current[1] = d.current[1];
  in.max_hsml = d.max_hsml;
  if (!(in.max_hsml > 0)) {
    double mx = 0;
    for (int i = 0; i < in.n; ++i) mx = std::max(mx, in.hsml[i]);
    in.max_hsml = mx * 10.;
  }
  in.g_nx = d.grid_nx; in.g_ny = d.grid_ny;
  in.g_x0 = d.grid_x0; in.g_y0 = d.grid_y0; in.g_dx = d.grid_dx; in.g_dy = d.grid_dy;
  in.g_u = d.grid_u; in.g_v = d.grid_v;
  return SIMP_OK;
}

int simp_batch_create(int precision, simp_batch **out) {
  if (!out) return SIMP_E_ARG;
  *out = nullptr;
  if (precision != SIMP_FP64 && precision != SIMP_MIXED) return SIMP_E_UNSUPPORTED;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SIMP_E_HIP;
  simp_batch *b = new (std::nothrow) simp_batch;
  if (!b) return SIMP_E_NOMEM;
  b->eng = new (std::nothrow) Engine(precision);
  if (!b->eng || !b->eng->stream()) {
    delete b->eng;
    delete b;
    return SIMP_E_HIP;
  }
  *out = b;
  return SIMP_OK;
}

int simp_batch_destroy(simp_batch *b) {
  if (!b) return SIMP_E_ARG;
  delete b->eng;
  delete b;
  return SIMP_OK;
}

int simp_batch_add(simp_batch *b, int count, const simp_instance_desc *desc, int *ids) {
  if (!b || count < 0 || (count > 0 && (!desc || !ids))) return SIMP_E_ARG;
  std::vector<InstanceInput> ins(count);
  for (int k = 0; k < count; ++k) {
    int rc = fill_from_desc(desc[k], ins[k]);
    if (rc) return rc;
  }
  std::vector<int> got;
  int rc = b->eng->add(ins, &got);
  if (rc) return map_rc(rc);
  for (int k = 0; k < count; ++k) ids[k] = got[k];
  return SIMP_OK;
}

int simp_batch_remove(simp_batch *b, int count, const int *ids) {
  if (!b || count < 0 || (count > 0 && !ids)) return SIMP_E_ARG;
  std::vector<int> v(ids, ids + count);
  for (int id : v)
    if (b->eng->index_of(id) < 0) return SIMP_E_ARG;
  return map_rc(b->eng->remove(v));
}

int simp_batch_update(simp_batch *b, int nremove, const int *remove_ids, int nadd,
                      const simp_instance_desc *desc, int *ids) {
  if (!b || nremove < 0 || nadd < 0 || (nremove > 0 && !remove_ids) ||
      (nadd > 0 && (!desc || !ids)))
    return SIMP_E_ARG;
  std::vector<int> rm(remove_ids, remove_ids + nremove);
  for (int id : rm)
    if (b->eng->index_of(id) < 0) return SIMP_E_ARG;
  std::vector<InstanceInput> ins(nadd);
  for (int k = 0; k < nadd; ++k) {
    int rc = fill_from_desc(desc[k], ins[k]);
    if (rc) return rc;
  }
  std::vector<int> got;
  int rc = b->eng->update(rm, ins, &got);
  if (rc) return map_rc(rc);
  for (int k = 0; k < nadd; ++k) ids[k] = got[k];
  return SIMP_OK;
}

int simp_batch_tune(simp_batch *b, int steps_per_trial, simp_tuning *chosen, double *us_start,
                    double *us_best) {
  if (!b || steps_per_trial < 1) return SIMP_E_ARG;
  if (b->eng->ninst() == 0) return SIMP_E_STATE;
  int rc = b->eng->autotune(steps_per_trial, us_start, us_best);
  if (rc == 0 && hipStreamSynchronize(b->eng->stream()) != hipSuccess) rc = -2;
  tuning_out(b->eng->tuning(), chosen);
  return map_rc(rc);
}

int simp_batch_set_tuning(simp_batch *b, const simp_tuning *t) {
  if (!b) return SIMP_E_ARG;
  simp::Tuning v;
  int rc = tuning_in(t, &v);
  if (rc) return rc;
  b->eng->set_tuning(v);
  return SIMP_OK;
}

int simp_batch_step(simp_batch *b, int nsteps, int sync) {
  if (!b || nsteps < 0) return SIMP_E_ARG;
  int rc = b->eng->steps(nsteps, true);
  if (rc) return map_rc(rc);
  if (sync && hipStreamSynchronize(b->eng->stream()) != hipSuccess) return SIMP_E_HIP;
  return SIMP_OK;
}

int simp_batch_sync(simp_batch *b) {
  if (!b) return SIMP_E_ARG;
  return hipStreamSynchronize(b->eng->stream()) == hipSuccess ? SIMP_OK : SIMP_E_HIP;
}

int simp_batch_status(simp_batch *b, int id, int *itimestep, int *steps_left, int *active,
                      int *status) {
  if (!b) return SIMP_E_ARG;
  const int k = b->eng->index_of(id);
  if (k < 0) return SIMP_E_ARG;
  if (hipStreamSynchronize(b->eng->stream()) != hipSuccess) return SIMP_E_HIP;
  int err = 0, act = 0;
  int rc = b->eng->download_status(k, itimestep, steps_left, &act, &err);
  if (rc) return map_rc(rc);
  if (active) *active = (act && !err) ? 1 : 0;   // a failed instance no longer steps
  if (status) {
    *status = SIMP_OK;
    if (err & simp::kErrCapacity) *status = SIMP_E_CAPACITY;
    if (err & simp::kErrNan) *status = SIMP_E_NAN;
    if (err & simp::kErrSync) *status = SIMP_E_HIP;
  }
  return SIMP_OK;
}

int simp_batch_download(simp_batch *b, int id, double *x, double *vx, double *rho, double *hsml,
                        double *hth, double *A, double *dt, double *current_time,
                        int *itimestep) {
  if (!b) return SIMP_E_ARG;
  const int k = b->eng->index_of(id);
  if (k < 0) return SIMP_E_ARG;
  int rc = b->eng->download_state(k, x, vx, rho, hsml, hth, A);
  if (rc) return map_rc(rc);
  return map_rc(b->eng->download_scalars(k, dt, current_time, itimestep, nullptr, nullptr,
                                         nullptr));
}

int simp_batch_interaction_stats(simp_batch *b, int id, int *maxip, int *minip, double *avgip,
                                 int *noip, long long *nip, int *itimestep) {
  if (!b) return SIMP_E_ARG;
  const int k = b->eng->index_of(id);
  if (k < 0) return SIMP_E_ARG;
  return stats_out(b->eng, k, b->eng->meta(k).n, maxip, minip, avgip, noip, nip, itimestep);
}

int simp_batch_count(simp_batch *b, int *instances, long long *particles) {
  if (!b) return SIMP_E_ARG;
  if (instances) *instances = b->eng->ninst();
  if (particles) *particles = b->eng->total_particles();
  return SIMP_OK;
}

int simp_batch_list(simp_batch *b, int cap, int *ids, int *itimestep, int *steps_left,
                    int *active, int *status) {
  if (!b || cap < 0) return -SIMP_E_ARG;
  std::vector<simp::InstDyn> dyn;
  int rc = b->eng->download_all_status(dyn);
  if (rc) return -map_rc(rc);
  const std::vector<int> &idv = b->eng->ids();
  const int n = (int)dyn.size();
  for (int k = 0; k < n && k < cap; ++k) {
    const simp::InstDyn &y = dyn[k];
    if (ids) ids[k] = idv[k];
    if (itimestep) itimestep[k] = y.itimestep;
    if (steps_left) steps_left[k] = y.steps_left;
    if (active) active[k] = (y.active && !y.err) ? 1 : 0;
    if (status) {
      status[k] = SIMP_OK;
      if (y.err & simp::kErrCapacity) status[k] = SIMP_E_CAPACITY;
      if (y.err & simp::kErrNan) status[k] = SIMP_E_NAN;
      if (y.err & simp::kErrSync) status[k] = SIMP_E_HIP;
    }
  }
  return n;
}

long long simp_instance_footprint(const simp_instance_desc *desc) {
  if (!desc) return -1;
  InstanceInput in;
  if (fill_from_desc(*desc, in) != SIMP_OK) return -1;
  return (long long)simp::instance_footprint(in);
}

}  // extern "C"
