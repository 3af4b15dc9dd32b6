"""Throughput and latency of the SPH viscous-plastic step on one MI355X.

Workload: the open-ice-edge case, 250,000 particles (500 x 500 lattice,
1 km spacing, 500 x 500 km, coast wall at x = 0, concentration ramp 1 -> 0.2
over the last 100 km, hsml = 1.3 s / sqrt(A), lognormal thickness,
Gaussian-random-field wind), synthetic, state resident in HBM.  A step is one
full predictor-corrector step (two right-hand-side evaluations + update).

Prints one JSON line: mixed-precision particle-steps/s (``value``), the FP64
rate, p50/p99 step latency, the autotuner's choice, the search statistics,
the dominant kernel's roofline (the kernel's own start / stop events on the library's stream) and HBM
traffic, the mixed-precision error per field, the per-call drop-in,
snapshots, the smaller box and vortex workloads, the packed ensemble and the
served request stream, and the host-CPU baseline (the compiled reference on
the same workload, bounded number of steps).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per particle of one launch of each heavy kernel
# (compulsory unique traffic, neighbour reads counted once, the per-stage
# figures of the step's byte budget below):
#   passA : x vx (32) mass hsml h A (32) -> sigma (3 x 8 | 3 x 4) + div (8)
#   passB : x vx mass h A (56) + sigma (24 | 12) + div (8) -> x vx h A (48),
#           the corrector also reading the step-start x vx h A (48) and
#           writing hsml (8); one launch of each phase per step, so the
#           per-launch figure is their mean
ALG_BYTES = {
    ("passA", 0): 64 + 24 + 8,
    ("passA", 1): 64 + 12 + 8,
    ("passB", 0): (136 + 192) / 2,
    ("passB", 1): (124 + 180) / 2,
}
# whole step: 576 B (FP64) / 512 B (mixed) per particle-step
STEP_BYTES = {0: 576, 1: 512}


def _done(section, t0):
    """A line on stderr when a section of the run has finished (the JSON
    result goes to stdout at the end)."""
    print(f"[bench] {section} done ({time.perf_counter() - t0:.1f} s)", file=sys.stderr,
          flush=True)


def _timed_steps(dev, steps):
    dev.sync()
    t0 = time.perf_counter()
    dev.step(steps, sync=True)
    return time.perf_counter() - t0


def _latencies(dev, count, eager=False):
    lat = []
    for _ in range(max(count, 1)):
        t0 = time.perf_counter()
        if eager:
            dev.profile_steps(1)
        else:
            dev.step(1, sync=True)
        lat.append(time.perf_counter() - t0)
    return np.array(lat)


def _cpu_threads():
    """Threads for the OpenMP run: the host's CPU share for one GPU (16 on
    the GPU hosts, where OMP_NUM_THREADS says so), else the CPUs this process
    may use."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _ref_rate(refrun, exe, sc, threads, target_s):
    """Particle-steps/s of one reference build on ``sc``: two runs of K0 and
    K0 + M steps, rate = n M / (wall(K0 + M) - wall(K0)).  The difference
    cancels what every run pays once -- reading the input and the one CSV
    snapshot the reference's time loop writes at the first step with
    current_time > 0 (time_integration.f90:30,61-65)."""
    def run(k):
        r = refrun.run(refrun.MODE_TIME_INTEGRATION, sc, integration_time=(k - 0.5) * sc.dt,
                       harness=exe, threads=threads)
        return r["itimestep"], r["wall"]
    k0, w0 = run(3)
    k1, w1 = run(8)
    per = max((w1 - w0) / max(k1 - k0, 1), 1e-4)
    m = int(max(5, min(2000, target_s / per)))
    k2, w2 = run(3 + m)
    steps, wall = k2 - k0, w2 - w0
    return {"threads": threads, "steps": steps, "seconds": round(wall, 3),
            "value": sc.n * steps / wall}


def cpu_baseline(sc_full, target_s=8.0):
    """The compiled reference on the host cores, on the full workload (the
    250k open ice edge with the reference's own hard-coded wind, VP time
    step): its OpenMP build on the host's CPU share and on one thread, and
    its serial build, all with maxn raised to 300,000.  ``value`` is the
    fastest of them."""
    from oracle import refrun
    sc = sc_full.copy()
    sc.wind_grid = None
    sc.wind = __import__("simp_amd.scenarios", fromlist=["x"]).REFERENCE_WIND
    sc.time = 0.0
    runs = {}
    if os.path.exists(refrun.HARNESS_OMP) and os.path.exists(refrun.HARNESS_BIG):
        T = _cpu_threads()
        runs[f"openmp_{T}_threads"] = _ref_rate(refrun, refrun.HARNESS_OMP, sc, T, target_s)
        runs["openmp_1_thread"] = _ref_rate(refrun, refrun.HARNESS_OMP, sc, 1, target_s)
        runs["serial"] = _ref_rate(refrun, refrun.HARNESS_BIG, sc, 1, target_s)
        best = max(runs, key=lambda k: runs[k]["value"])
        b = runs[best]
        return {"value": b["value"], "unit": "particle-steps/s", "cores": b["threads"],
                "kind": "reference", "best": best, "runs": runs,
                "nproc": os.cpu_count(), "threads_available": _cpu_threads(),
                "sample": f"reference Fortran (amdflang -O2, maxn=300000) on the full "
                          f"workload: {sc.n} particles, {sc.nb} boundary particles, the "
                          f"reference's own wind; rate from the difference of a 3-step and a "
                          f"{3 + b['steps']}-step run ({b['seconds']:.2f} s), best of the "
                          f"OpenMP build on {_cpu_threads()} threads, on 1 thread and the "
                          f"serial build"}
    # no compiled reference on this host: the C restatement, serial
    from oracle import capi
    o = capi.Oracle()
    s = sc.copy()
    t0 = time.perf_counter()
    o.run_pc(s, 2)
    per_step = (time.perf_counter() - t0) / 2
    steps = int(max(3, min(400, target_s / max(per_step, 1e-6))))
    t0 = time.perf_counter()
    o.run_pc(s, steps)
    wall = time.perf_counter() - t0
    return {"value": sc.n * steps / wall, "unit": "particle-steps/s", "cores": 1,
            "kind": "port",
            "sample": f"C restatement of the reference (serial) on the full workload: "
                      f"{sc.n} particles, {steps} steps, {wall:.2f} s"}


def _copy_calibration(mib):
    """A plain device-to-device copy of known size through the HIP runtime
    (its blit kernel is the reference for the FETCH_SIZE / WRITE_SIZE
    counters)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    nbytes = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0   # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def hbm_traffic(kernel_substr, timeout=240):
    """HBM bytes per launch of the kernel whose name contains kernel_substr,
    from two rocprofv3 counter passes (FETCH_SIZE, then WRITE_SIZE) over a
    short mixed-precision run of this script (eager launches: the counters
    of a kernel do not depend on how it was launched, and the profiler's
    instrumentation of graph launches is the fragile part).  Each counter is scaled by
    what it reads for plain device-to-device copies of known size (3 x 256
    MiB, run first in the profiled process).  Returns (bytes, factors) or
    (None, reason)."""
    import csv
    import shutil
    import subprocess
    import tempfile
    exe = shutil.which("rocprofv3")
    if exe is None:
        return None, "rocprofv3 not found"
    expect_kib = 3 * 256 * 1024
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            d = os.path.join(tmp, counter)
            cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "c", "--",
                   sys.executable, os.path.abspath(__file__), "--eager", "--precision",
                   "mixed", "--steps", "10", "--warmup", "3", "--latency-steps", "1",
                   "--no-cpu", "--copy-calib", "256"]
            try:
                subprocess.run(cmd, check=True, timeout=timeout, cwd=tmp,
                               stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
            except Exception as e:  # noqa: BLE001
                return None, f"rocprofv3 {counter} pass failed: {e}"
            rows = []
            for root, _, files in os.walk(d):
                for f in files:
                    if f.endswith("counter_collection.csv"):
                        rows += list(csv.DictReader(open(os.path.join(root, f))))
            rows.sort(key=lambda r: int(r["Dispatch_Id"]))
            calib, vals = 0.0, []
            seen_simp = False
            for r in rows:
                name = r["Kernel_Name"]
                if "simp::" in name:
                    seen_simp = True
                if not seen_simp and "copyBuffer" in name:
                    calib += float(r["Counter_Value"])
                if kernel_substr in name:
                    vals.append(float(r["Counter_Value"]))
            if calib <= 0 or not vals:
                return None, f"no {counter} samples"
            out[counter] = (sum(vals) / len(vals), expect_kib / calib)
    fetch, ff = out["FETCH_SIZE"]
    write, fw = out["WRITE_SIZE"]
    return 1024.0 * (fetch * ff + write * fw), {"fetch_factor": ff, "write_factor": fw,
                                                 "fetch_kib": fetch, "write_kib": write}


def bench_forced_search(sc, steps):
    """The same step with the cell lists and candidate lists rebuilt at every
    evaluation, as the reference's bucket_search searches (same results)."""
    from simp_amd import _lib
    from simp_amd.api import Simp
    out = {"note": "cell lists and candidate lists rebuilt at every evaluation "
                   "(simp_set_search(h, 1)); same results as the headline",
           "steps": steps}
    for name, prec in (("mixed", _lib.SIMP_MIXED), ("fp64", _lib.SIMP_FP64)):
        dev = Simp.for_scenario(sc, precision=prec)
        dev.set_search(True)
        dev.upload(sc)
        dev.step(10, sync=True)
        wall = _timed_steps(dev, steps)
        st = dev.search_stats()
        assert st["cell_rebuilds"] == st["evaluations"]
        out[name] = {"value": sc.n * steps / wall, "ms_per_step": 1e3 * wall / steps}
        dev.close()
    return out


def bench_field_error(sc):
    """Mixed-precision error per field against the FP64 path (which the
    parity tests pin to the oracle) at the workloads' own step counts: the
    open ice edge (250k particles, 1,000 steps) and the Rankine vortex (40k
    particles, 500 steps).  norm = max|a - b| / max|b|; elem = largest
    |a_i - b_i| / (|b_i| + 1e-3 max|b|); delta = normwise error of the change
    over the run (rho, vx, h, A); disp = displacement x - x0."""
    from simp_amd import accuracy
    from simp_amd import scenarios as S

    def rounded(tab):
        return {k: {m: float(f"{v:.3e}") for m, v in r.items() if m in ("norm", "elem", "delta")}
                for k, r in tab.items()}
    out = {"reference": "FP64 GPU path", "target_norm": 1e-4}
    t_edge = accuracy.report(sc, 1000)
    t_vort = accuracy.report(S.vortex(), 500)
    out["ice_edge_250k_1000_steps"] = rounded(t_edge)
    out["vortex_40k_500_steps"] = rounded(t_vort)
    out["worst_norm"] = max(accuracy.worst(t_edge), accuracy.worst(t_vort))
    return out


def bench_dropin(sc, calls=10):
    """The per-call drop-in (simp_single_step): one right-hand side on host
    arrays per call, the state copied in and the tendencies and diagnostics
    copied out over PCIe, the way a Fortran time_integration calls it."""
    from simp_amd import _lib
    from simp_amd.api import Simp
    dev = Simp.for_scenario(sc, precision=_lib.SIMP_FP64)
    out = dev.single_step(sc)
    ts = []
    for _ in range(calls):
        t0 = time.perf_counter()
        dev.single_step(sc, out=out)
        ts.append(time.perf_counter() - t0)
    dev.close()
    dt = sum(ts)
    mb = (sc.n * 9 * 8 + sc.n * 15 * 8) / 1e6
    return {"calls": calls, "ms_per_call": 1e3 * dt / calls,
            "ms_per_call_min": 1e3 * min(ts),
            "particle_evaluations_per_s": sc.n * calls / dt,
            "pcie_mb_per_call": mb,
            "note": "PCIe-inclusive, FP64; one call = one right-hand side (half a step): "
                    "the state copied in, NNPS and the sums on the GPU, tendencies and "
                    "diagnostics copied out; the resident layout is reused between calls"}


def bench_ensemble(steps=500, baseline_steps=50):
    """Config 4: 64 closed boxes, particle counts lognormal in [2k, 60k],
    box 50-300 km, random uniform wind, 500 steps; all stepped in one packed
    launch per kernel (mixed precision), against the plain FP64 path
    stepping each instance with its own launches (one instance after the
    other)."""
    from simp_amd import _lib
    from simp_amd import scenarios as S
    from simp_amd.api import Simp
    from simp_amd.ensemble import Ensemble
    scs = S.ensemble(count=64)
    total = sum(sc.n for sc in scs)
    out = {"instances": len(scs), "particles": total,
           "n_min": min(sc.n for sc in scs), "n_max": max(sc.n for sc in scs)}
    for prec, key in ((_lib.SIMP_MIXED, "packed_mixed"), (_lib.SIMP_FP64, "packed_fp64")):
        ens = Ensemble(prec)
        ens.add(scs)
        ens.step(5)
        out[key + "_tuning"] = ens.tune(10)
        t0 = time.perf_counter()
        ens.step(steps)
        dt = time.perf_counter() - t0
        out[key] = total * steps / dt
        out[key + "_ms_per_step"] = 1e3 * dt / steps
        ens.close()
    hs = []
    for sc in scs:
        h = Simp.for_scenario(sc, precision=_lib.SIMP_FP64)
        h.upload(sc)
        h.step(2)
        hs.append(h)
    t0 = time.perf_counter()
    for h in hs:
        h.step(baseline_steps, sync=True)
    dt = time.perf_counter() - t0
    out["per_instance_fp64"] = total * baseline_steps / dt
    out["packed_mixed_over_per_instance_fp64"] = out["packed_mixed"] / out["per_instance_fp64"]
    for h in hs:
        h.close()
    return out


def bench_snapshots(sc, steps=10):
    """Snapshots of the 250k run after a few device-resident steps: one CSV
    snapshot in the reference's format (the host writer, to a temporary
    folder) and one exact snapshot saved and restored."""
    import tempfile
    from simp_amd import _lib, output
    from simp_amd.api import Simp
    sim = Simp.for_scenario(sc, precision=_lib.SIMP_MIXED)
    sim.upload(sc)
    sim.set_diagnostics(True)
    sim.step(steps)
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        output.create(tmp)
        t0 = time.perf_counter()
        output.snapshot(sim, tmp, sc.mass, sc.nb)
        out["csv_ms"] = 1e3 * (time.perf_counter() - t0)
        size = sum(os.path.getsize(os.path.join(tmp, f)) for f in os.listdir(tmp))
    out["csv_bytes"] = size
    out["csv_mb_per_s"] = size / 1e6 / (out["csv_ms"] / 1e3)
    t0 = time.perf_counter()
    blob = sim.snapshot()
    out["exact_save_ms"] = 1e3 * (time.perf_counter() - t0)
    out["exact_bytes"] = len(blob)
    other = Simp.for_scenario(sc, precision=_lib.SIMP_MIXED)
    t0 = time.perf_counter()
    other.restore(blob)
    other.sync()
    out["exact_restore_ms"] = 1e3 * (time.perf_counter() - t0)
    out["note"] = ("CSV: download + host formatting + write (includes PCIe); "
                   "exact: state, masses, boundary, kernel_zero and hsml_ij store")
    sim.close()
    other.close()
    return out


def bench_configs(cpu=True):
    """The two smaller single-simulation workloads at their own sizes and
    step counts: the closed box (2,500 particles, 200 steps) and the Rankine
    vortex (40,000 particles, 500 steps), each in FP64 and mixed precision
    with p50/p99 step latency, and the compiled reference on the host cores
    (OpenMP build, bounded sample) beside them."""
    from simp_amd import _lib
    from simp_amd import scenarios as S
    from simp_amd.api import Simp
    out = {}
    for key, sc, steps in (("box_2500_200_steps", S.box(), 200),
                           ("vortex_40k_500_steps", S.vortex(), 500)):
        r = {"particles": sc.n, "boundary_particles": sc.nb, "steps": steps}
        for name, prec in (("fp64", _lib.SIMP_FP64), ("mixed", _lib.SIMP_MIXED)):
            dev = Simp.for_scenario(sc, precision=prec)
            dev.upload(sc)
            dev.step(10, sync=True)
            dev.tune(10)
            dev.upload(sc)
            wall = _timed_steps(dev, steps)
            lat = _latencies(dev, 100, False)
            r[name] = {"value": sc.n * steps / wall, "ms_per_step": 1e3 * wall / steps,
                       "p50_step_ms": 1e3 * float(np.percentile(lat, 50)),
                       "p99_step_ms": 1e3 * float(np.percentile(lat, 99))}
            dev.close()
        if cpu:
            from oracle import refrun
            if os.path.exists(refrun.HARNESS_OMP):
                c = sc.copy()
                c.wind_grid = None
                c.wind = S.REFERENCE_WIND
                c.time = 0.0
                T = _cpu_threads()
                best = max((_ref_rate(refrun, refrun.HARNESS_OMP, c, t, 2.0) for t in (T, 1)),
                           key=lambda v: v["value"])
                r["cpu_reference"] = best
                r["gpu_fp64_over_cpu"] = r["fp64"]["value"] / best["value"]
        out[key] = r
    return out


def bench_serve(count=80, budget_gib=16.0, chunk=10):
    """Config 5: Poisson arrivals (20/s) of closed-box simulations, N in
    [5k, 250k], 100-2000 steps each, HBM-budget admission, continuous
    batching, mixed precision."""
    from simp_amd import _lib
    from simp_amd.serve import Server, poisson_requests
    reqs = poisson_requests(count=count)
    srv = Server(precision=_lib.SIMP_MIXED, budget_bytes=int(budget_gib * (1 << 30)),
                 chunk=chunk)
    m = srv.run(reqs)
    m.pop("results")
    m["budget_gib"] = budget_gib
    m["chunk_steps"] = chunk
    return m


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--latency-steps", type=int, default=300)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="first run a plain device-to-device copy of MIB MiB (the "
                         "reference kernel for scaling HBM counters)")
    ap.add_argument("--precision", choices=("both", "mixed", "fp64"), default="both")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 counter passes for roofline.traffic")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default tuning instead of running the autotuner")
    ap.add_argument("--eager", action="store_true",
                    help="launch the step's kernels directly instead of replaying the captured "
                         "step graph (for kernel tracing; slower)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the 2,500-particle box and 40k vortex workloads")
    ap.add_argument("--workload", choices=("ice_edge", "ensemble", "serve"), default="ice_edge",
                    help="ensemble / serve: print only that workload's line")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the SPH-VP step of one simulation (or one "
              "packed ensemble) runs on a single MI355X here; run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    from simp_amd import _lib
    from simp_amd import scenarios as S
    from simp_amd.api import Simp

    if args.copy_calib > 0:
        _copy_calibration(args.copy_calib)
    if args.workload == "ensemble":
        print(json.dumps({"workload": "ensemble", **bench_ensemble(args.steps)}), flush=True)
        return
    if args.workload == "serve":
        print(json.dumps({"workload": "serve", **bench_serve()}), flush=True)
        return
    sc = S.ice_edge()
    n = sc.n
    res = {}
    prof = {}
    tuned = {}
    precs = {"both": (_lib.SIMP_FP64, _lib.SIMP_MIXED), "mixed": (_lib.SIMP_MIXED,),
             "fp64": (_lib.SIMP_FP64,)}[args.precision]
    for prec in precs:
        t_start = time.perf_counter()
        dev = Simp.for_scenario(sc, precision=prec)
        dev.upload(sc)
        if args.eager:
            # every launch of the step direct (timed with events), no captured
            # graph: for rocprofv3's kernel tracing, which fails on the
            # graph launches of this build (host segfault in the tracer)
            dev.profile_steps(args.warmup)
            t0 = time.perf_counter()
            dev.profile_steps(args.steps)
            wall = time.perf_counter() - t0
        else:
            dev.step(args.warmup, sync=True)
            # the autotuner picks the kernels' batch sizes and the list skin
            # on this workload (its trial steps are steps of the simulation)
            tuned[prec] = dev.tune(30) if not args.no_tune else dev.get_tuning()
            wall = _timed_steps(dev, args.steps)
        lat = _latencies(dev, args.latency_steps, args.eager)
        res[prec] = {"rate": n * args.steps / wall, "ms": 1e3 * wall / args.steps,
                     "p50": 1e3 * float(np.percentile(lat, 50)),
                     "p99": 1e3 * float(np.percentile(lat, 99)),
                     "search": dev.search_stats()}
        prof[prec] = dev.profile_steps(20)
        dev.close()
        _done("headline, " + ("mixed" if prec == _lib.SIMP_MIXED else "fp64"), t_start)

    if args.precision != "both":
        r = res[precs[0]]
        print(json.dumps({"precision": args.precision, "particle_steps_per_s": r["rate"],
                          "ms_per_step": r["ms"], "kernels_ms": {
                              k: v[0] for k, v in prof[precs[0]].items()}}), flush=True)
        return
    mixed = res[_lib.SIMP_MIXED]
    fp64 = res[_lib.SIMP_FP64]
    # dominant kernel of the mixed path
    pm = prof[_lib.SIMP_MIXED]
    name = max(("passA", "passB"), key=lambda k: pm[k][0])
    kms = pm[name][0]
    alg = ALG_BYTES[(name, 1)] * n
    achieved = alg / (kms * 1e-3) / 1e9
    line = {
        "metric": "particle-steps/sec at 250k particles (FP32-mixed vs FP64) + p50/p99 step latency",
        "value": mixed["rate"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-pair-sums+f64",
        "data": "synthetic",
        "config": {"workload": "ice_edge: 250,000 particles, 500x500 km, coast wall, "
                               "concentration ramp, variable hsml, GRF wind; mixed precision",
                   "particles": n, "boundary_particles": sc.nb, "time_step": "VP (tse=2)"},
        "fp64": {"value": fp64["rate"], "ms_per_step": fp64["ms"],
                 "p50_step_ms": fp64["p50"], "p99_step_ms": fp64["p99"]},
        "p50_step_ms": mixed["p50"],
        "p99_step_ms": mixed["p99"],
        "mixed_over_fp64": mixed["rate"] / fp64["rate"],
        "step_hbm_roofline_frac": {
            "mixed": STEP_BYTES[1] * mixed["rate"] / (HBM_PEAK_GBS * 1e9),
            "fp64": STEP_BYTES[0] * fp64["rate"] / (HBM_PEAK_GBS * 1e9)},
        "kernels_ms": {k: round(v[0], 5) for k, v in pm.items()},
        "roofline": {"bound": "hbm", "kernel": f"k_passB<mixed>" if name == "passB"
                     else "k_passA<mixed>",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "alg_bytes_per_launch": alg, "avg_launch_ms": kms},
    }
    if tuned:
        line["autotune"] = {
            "note": "chosen by simp_tune on this workload after the warm-up (pass A / pass B "
                    "neighbour batches, candidate-list skin as a fraction of MINVAL(hsml)); "
                    "us_start: the default setting, us_best: the chosen one",
            **{name: tuned[p] for name, p in (("mixed", _lib.SIMP_MIXED), ("fp64", _lib.SIMP_FP64))
               if p in tuned}}
    line["search"] = {
        "note": "evaluations since the upload (warm-up, timed and latency steps) and the "
                "share that rebuilt the cell lists / the candidate lists",
        **{name: {"evaluations": r["search"]["evaluations"],
                  "cell_rebuild_rate": r["search"]["cell_rebuilds"] / r["search"]["evaluations"],
                  "candidate_rebuild_rate":
                      r["search"]["candidate_rebuilds"] / r["search"]["evaluations"]}
           for name, r in (("mixed", mixed), ("fp64", fp64))}}
    if args.eager:
        # (the measurements below replay captured graphs)
        line["launch"] = "eager (kernel tracing); ms_per_step includes the event records"
        print(json.dumps(line), flush=True)
        return
    t0 = time.perf_counter()
    line["search_every_evaluation"] = bench_forced_search(sc, min(args.steps, 300))
    _done("search at every evaluation", t0)
    t0 = time.perf_counter()
    line["field_error"] = bench_field_error(sc)
    _done("field error", t0)
    t0 = time.perf_counter()
    line["dropin_single_step"] = bench_dropin(sc)
    line["snapshots"] = bench_snapshots(sc)
    _done("drop-in and snapshots", t0)
    if not args.no_configs:
        t0 = time.perf_counter()
        line["configs"] = bench_configs(cpu=not args.no_cpu)
        _done("box and vortex workloads", t0)
    if not args.no_ensemble:
        t0 = time.perf_counter()
        line["ensemble"] = bench_ensemble()
        _done("ensemble", t0)
    if not args.no_serve:
        t0 = time.perf_counter()
        line["serve"] = bench_serve()
        _done("serve", t0)
    if not args.no_traffic:
        # (pass B: both phase kernels, k_passB<true, false, 0|1>)
        kname = "k_passB<true, false," if name == "passB" else "k_passA<true, false,"
        t0 = time.perf_counter()
        traffic, info = hbm_traffic(kname)
        _done("HBM traffic counters", t0)
        line["roofline"]["traffic"] = traffic
        line["roofline"]["traffic_note"] = info
    if not args.no_cpu:
        t0 = time.perf_counter()
        cb = cpu_baseline(sc)
        _done("CPU baseline", t0)
        cb["gpu_fp64_over_cpu"] = fp64["rate"] / cb["value"]
        cb["gpu_mixed_over_cpu"] = mixed["rate"] / cb["value"]
        line["cpu_baseline"] = cb
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
