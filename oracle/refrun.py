"""Driver of the compiled reference (``_ref/simp_ref_harness``).

TEST INFRASTRUCTURE ONLY.  The harness is built from the reference's own
Fortran sources by ``build_ref.sh``; this module writes its binary input,
runs it in a scratch directory and parses its FP64 output.  Used to write
the golden vectors under ``tests/golden`` and as the CPU baseline.
"""
from __future__ import annotations

import os
import resource
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(HERE, "_ref", "simp_ref_harness")
#: the same harness with single_step replaced by the GPU shim
#: (simp_amd/fortran/single_step_gpu.f90 -> libsimp_amd.so)
HARNESS_GPU = os.path.join(HERE, "_ref", "simp_ref_harness_gpu")
#: the harness built with summation density (den_alg = 1)
HARNESS_DEN1 = os.path.join(HERE, "_ref", "simp_ref_harness_den1")
HARNESS_GPU_DEN1 = os.path.join(HERE, "_ref", "simp_ref_harness_gpu_den1")
#: the harness built with the stress_2 particle approximation (pa_sph = 2)
HARNESS_PA2 = os.path.join(HERE, "_ref", "simp_ref_harness_pa2")
HARNESS_GPU_PA2 = os.path.join(HERE, "_ref", "simp_ref_harness_gpu_pa2")
#: the harnesses built with the Hibler (eos_alg = 1) and Gutfraind (2) pressure
HARNESS_EOS1 = os.path.join(HERE, "_ref", "simp_ref_harness_eos1")
HARNESS_EOS2 = os.path.join(HERE, "_ref", "simp_ref_harness_eos2")
HARNESS_GPU_EOS1 = os.path.join(HERE, "_ref", "simp_ref_harness_gpu_eos1")
HARNESS_GPU_EOS2 = os.path.join(HERE, "_ref", "simp_ref_harness_gpu_eos2")
#: the harnesses built with Mohr-Coulomb rheology (rheology = 2), the
#: Wendland C2 / C4 kernels (skf = 4, 5) and the Euler / leap-frog schemes
#: (int_algorithm = 1, 3), and their GPU drop-in counterparts
HARNESS_MC = os.path.join(HERE, "_ref", "simp_ref_harness_mc")
HARNESS_C2 = os.path.join(HERE, "_ref", "simp_ref_harness_c2")
HARNESS_C4 = os.path.join(HERE, "_ref", "simp_ref_harness_c4")
HARNESS_EULER = os.path.join(HERE, "_ref", "simp_ref_harness_euler")
#: the harnesses built with skf = 1, 2, 3, 7, 8
HARNESS_SKF = {k: os.path.join(HERE, "_ref", f"simp_ref_harness_k{k}") for k in (1, 2, 3, 7, 8)}
HARNESS_GPU_SKF = {k: os.path.join(HERE, "_ref", f"simp_ref_harness_gpu_k{k}")
                   for k in (1, 2, 3, 7, 8)}
HARNESS_LEAP = os.path.join(HERE, "_ref", "simp_ref_harness_leap")
HARNESS_GPU_MC = os.path.join(HERE, "_ref", "simp_ref_harness_gpu_mc")
HARNESS_GPU_C2 = os.path.join(HERE, "_ref", "simp_ref_harness_gpu_c2")
HARNESS_GPU_C4 = os.path.join(HERE, "_ref", "simp_ref_harness_gpu_c4")
HARNESS_GPU_EULER = os.path.join(HERE, "_ref", "simp_ref_harness_gpu_euler")
HARNESS_GPU_LEAP = os.path.join(HERE, "_ref", "simp_ref_harness_gpu_leap")
#: the harness built with the Monaghan time step (tse = 1)
HARNESS_TSE1 = os.path.join(HERE, "_ref", "simp_ref_harness_tse1")
HARNESS_GPU_TSE1 = os.path.join(HERE, "_ref", "simp_ref_harness_gpu_tse1")
#: the harness built with continuity density (den_alg = 2)
HARNESS_DEN2 = os.path.join(HERE, "_ref", "simp_ref_harness_den2")
HARNESS_GPU_DEN2 = os.path.join(HERE, "_ref", "simp_ref_harness_gpu_den2")
#: the harnesses built with sle = 0 / 1 (also with Euler, with continuity
#: density), XSPH, the two artificial viscosities (Monaghan's also with the
#: Monaghan time step), the Wang / Marquis boundary forces, boundary
#: friction (with six kernels, with Wang's and Marquis' forces), the
#: boundary pressure off, wind, water drag and pressure off, and viscosity
#: off (with pressure off; with the Hibler eos), by tag
VARIANT_TAGS = ("sle0", "sle1", "sle1eu", "sle1d2", "xsph", "avhk", "avmon", "avtse1",
                "bfa2", "bfa3", "fric", "fric1", "fricg", "fricq", "fric4", "fric5", "bfa2f",
                "bfa3f", "nobp", "noforce", "novisc", "novisc1")
HARNESS_VARIANT = {t: os.path.join(HERE, "_ref", f"simp_ref_harness_{t}") for t in VARIANT_TAGS}
HARNESS_GPU_VARIANT = {t: os.path.join(HERE, "_ref", f"simp_ref_harness_gpu_{t}")
                       for t in VARIANT_TAGS}
#: the OpenMP build sized for the 250k workload (maxn = 300000), timings only
HARNESS_OMP = os.path.join(HERE, "_ref", "simp_ref_harness_omp")
#: the serial build with the same capacity
HARNESS_BIG = os.path.join(HERE, "_ref", "simp_ref_harness_big")
MAXN = 80000    # globals.f90:22
MAXNB = 5000    # globals.f90:23
MAXN_OMP = 300000

MODE_NNPS, MODE_SINGLE_STEP, MODE_TIME_INTEGRATION, MODE_TWO_STATES = 1, 2, 3, 4


def available() -> bool:
    return os.path.exists(HARNESS)


def _write_state(f, sc):
    for a in (sc.x, sc.vx, sc.mass, sc.rho, sc.hsml, sc.A, sc.h):
        f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
    if sc.nb:
        f.write(np.ascontiguousarray(sc.xb, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(sc.massb, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(sc.b_normal(), dtype="<f8").tobytes())


def _write_input(path, mode, sc, reps, integration_time, second=None):
    with open(path, "wb") as f:
        f.write(struct.pack("<4i", mode, sc.n, sc.nb, reps))
        f.write(struct.pack("<4d", sc.dt, sc.time, integration_time, sc.max_hsml))
        _write_state(f, sc)
        if second is not None:
            assert second.n == sc.n and second.nb == sc.nb
            _write_state(f, second)


def _read_output(path, n):
    buf = open(path, "rb").read()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(buf, dtype=dtype, count=count, offset=off).copy()
        off += a.nbytes
        return a

    nip, bnip, itimestep, _ = take("<i4", 4)
    dt, t, wall = take("<f8", 3)
    out = {"nip": int(nip), "bnip": int(bnip), "itimestep": int(itimestep),
           "dt": float(dt), "time": float(t), "wall": float(wall)}
    out["p_i"] = take("<i4", nip); out["p_j"] = take("<i4", nip)
    out["r_ij"] = take("<f8", nip); out["hsml_ij"] = take("<f8", nip)
    out["bp_i"] = take("<i4", bnip); out["bp_j"] = take("<i4", bnip)
    out["br_ij"] = take("<f8", bnip)
    out["x"] = take("<f8", 2 * n).reshape(n, 2)
    out["vx"] = take("<f8", 2 * n).reshape(n, 2)
    for k in ("rho", "hsml", "A", "h"):
        out[k] = take("<f8", n)
    out["dxdt"] = take("<f8", 2 * n).reshape(n, 2)
    out["dvxdt"] = take("<f8", 2 * n).reshape(n, 2)
    out["dhdt"] = take("<f8", n); out["dAdt"] = take("<f8", n)
    e11, e12, e22 = take("<f8", n), take("<f8", n), take("<f8", n)
    out["eps"] = np.stack([e11, e12, e22], axis=1)
    out["zeta"] = take("<f8", n); out["eta"] = take("<f8", n); out["p"] = take("<f8", n)
    s11, s12, s22 = take("<f8", n), take("<f8", n), take("<f8", n)
    out["sigma"] = np.stack([s11, s12, s22], axis=1)
    out["w"] = take("<f8", nip)
    out["kernel_zero"] = take("<f8", n)
    return out


def _big_stack():
    # the reference keeps its per-call work arrays (single_step.f90:31,
    # time_integration.f90:190-194, nnps.f90:196) on the stack
    try:
        resource.setrlimit(resource.RLIMIT_STACK,
                           (resource.RLIM_INFINITY, resource.RLIM_INFINITY))
    except (ValueError, OSError):
        soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
        resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))


def run(mode, sc, reps=1, integration_time=0.0, threads=1, timeout=3600, second=None,
        harness=None):
    """Run the reference harness on scenario ``sc`` (and ``second`` for
    MODE_TWO_STATES); returns a dict of its outputs.  The reference's wind is
    its own hard-coded field."""
    exe = harness or HARNESS
    if not os.path.exists(exe):
        raise FileNotFoundError(exe)
    maxn = MAXN_OMP if exe in (HARNESS_OMP, HARNESS_BIG) else MAXN
    if sc.n > maxn or sc.nb > MAXNB:
        raise ValueError("scenario exceeds the reference's static capacity")
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        _write_input(fin, mode, sc, reps, integration_time, second)
        env = dict(os.environ, PWD=d, OMP_NUM_THREADS=str(threads), OMP_STACKSIZE="256M")
        subprocess.run([exe, fin, fout], cwd=d, env=env, check=True,
                       stdout=subprocess.DEVNULL, timeout=timeout,
                       preexec_fn=_big_stack)
        return _read_output(fout, sc.n)
