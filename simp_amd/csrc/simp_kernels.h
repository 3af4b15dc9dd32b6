// simp_kernels.h -- host-side launchers of the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "simp_device.h"

namespace simp {

// Start / end events of each kernel launched while a pointer to this is set
// in g_kernel_events (the calling thread's launches only); the owner
// destroys them.
struct KernelEvents {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
};
extern thread_local KernelEvents *g_kernel_events;

void launch_partial(const DevState &d, int buf, hipStream_t st);
void launch_finalize(const DevState &d, int end_of_step, hipStream_t st);
// end of an evaluation outside a step (single evaluation, pair export)
void launch_fold(const DevState &d, hipStream_t st);
// every instance collects interaction statistics in the next evaluation
void launch_stats_reset(const DevState &d, hipStream_t st);
void launch_count(const DevState &d, int buf, hipStream_t st);
// gate: device flag; the scan does nothing while it reads 0 (nullptr: always runs)
void launch_scan(int *in, int L, int *tmp, int *out, int zero_in, const int *gate,
                 hipStream_t st);
// the rebuild of the cell lists of the instances with a changed key (counts
// are current): slots, cell starts, new order, state move -- one launch, a
// flag read between rebuilds.  tmp: the scan's block sums
void launch_rebuild(const DevState &d, int buf, int *tmp, hipStream_t st);
// ba / bb: neighbours whose loads go out together in pass A / pass B (tuning
// parameters; 0 or an untuned value: the default)
void launch_passA(const DevState &d, int buf, int mixed, int xsw, int want_diag, int ba,
                  hipStream_t st);
void launch_pairtot(const DevState &d, hipStream_t st);
void launch_ranks(const DevState &d, int buf, hipStream_t st);
// phase: 0 predictor, 1 corrector, 2 single evaluation, 3 Euler evaluation
// (then launch_euler), 4 second leap-frog stage
void launch_passB(const DevState &d, int mixed, int xsw, int phase, int bb, hipStream_t st);
void launch_euler(const DevState &d, hipStream_t st);
// continuity density after an update (instances with den_alg = 2): mode 0
// after the predictor, 1 before the Euler update, 2 after the leap-frog's
// second evaluation
void launch_condens(const DevState &d, int mode, hipStream_t st);
// artificial viscosity (instances with avis != 0), after pass A and the pair
// ranks (launch_pairtot, launch_scan): each pair's pi_ij in the packed pair
// ranks, the value a pair that does not set one keeps from the pair before
// it, and every list entry's pair rank
void launch_av(const DevState &d, int buf, hipStream_t st);
// end-of-step pair pass (instances with endp): mode 0 after the corrector or
// the leap-frog's second evaluation, 1 after the Euler evaluation, before
// launch_euler
void launch_endpass(const DevState &d, int mode, hipStream_t st);
void launch_export_pairs(const DevState &d, int buf, int cap, int *pi, int *pj, double *rij,
                         double *hij, double *w, double *dwdx, hipStream_t st);
void launch_export_bpairs(const DevState &d, int buf, int nb, int bcap, int *bpi, int *bpj,
                          double *brij, hipStream_t st);
// one instance laid out in caller memory: its block table (d.blk, d.nblk),
// its state from the caller's (2, n) arrays staged on the device, and the
// results of a single evaluation into the caller's order (15 n doubles)
void launch_single_blocks(const DevState &d, int n, hipStream_t st);
void launch_single_unpack(const DevState &d, const double *x2, const double *v2,
                          const double *xb2, int n, int nb, hipStream_t st);
void launch_single_pack(const DevState &d, double *out, int n, int want_t, int want_d,
                        hipStream_t st);

}  // namespace simp
