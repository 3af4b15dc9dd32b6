"""The drop-in: the reference's own time_integration (compiled from its
sources, unchanged) driving the GPU single_step through the iso_c_binding
shim, against the same build with the reference's single_step.  Both
binaries are prebuilt under oracle/_ref by oracle/build_ref.sh."""
from __future__ import annotations

import os

import numpy as np
import pytest

from golden_io import load, scenario
from oracle import refrun

pytestmark = pytest.mark.gpu


def _input_of(tag):
    """The golden whose input state a variant's drop-in case runs (both
    binaries run it; the comparison is between them): the Wang force
    survives two steps; the friction variants of other kernels take the
    ice edge of the C6 friction case."""
    if tag in ("bfa2", "bfa2f"):
        return f"edge3600_pc2_{tag}.npz"
    if tag.startswith("fric"):
        return "edge3600_pc10_fric.npz"
    return f"edge3600_pc10_{tag}.npz"


CASES = [("box2500_pc20.npz", refrun.HARNESS, refrun.HARNESS_GPU),
         ("edge3600_pc10.npz", refrun.HARNESS, refrun.HARNESS_GPU),
         # summation density: the reference's sum_density runs on the host on
         # the kernel values and kernel_zero the shim exports
         ("edge3600_pc10_den1.npz", refrun.HARNESS_DEN1, refrun.HARNESS_GPU_DEN1),
         # stress_2 (pa_sph = 2): the shim passes the module's pa_sph
         ("edge3600_pc10_pa2.npz", refrun.HARNESS_PA2, refrun.HARNESS_GPU_PA2),
         # Gutfraind pressure (eos_alg = 2)
         ("edge3600_pc10_eos1.npz", refrun.HARNESS_EOS1, refrun.HARNESS_GPU_EOS1),
         ("edge3600_pc10_eos2.npz", refrun.HARNESS_EOS2, refrun.HARNESS_GPU_EOS2),
         # Mohr-Coulomb rheology, Wendland C2 / C4 kernels: the shim passes the
         # module's rheology and skf
         ("edge3600_pc10_mc.npz", refrun.HARNESS_MC, refrun.HARNESS_GPU_MC),
         ("edge3600_pc10_c2.npz", refrun.HARNESS_C2, refrun.HARNESS_GPU_C2),
         ("edge3600_pc10_c4.npz", refrun.HARNESS_C4, refrun.HARNESS_GPU_C4),
         # the reference's own Euler and leap-frog loops calling the GPU
         # single_step
         ("edge3600_pc10_euler.npz", refrun.HARNESS_EULER, refrun.HARNESS_GPU_EULER),
         ("edge3600_pc10_leap.npz", refrun.HARNESS_LEAP, refrun.HARNESS_GPU_LEAP),
         ("edge3600_pc10_tse1.npz", refrun.HARNESS_TSE1, refrun.HARNESS_GPU_TSE1),
         ("edge3600_pc10_den2.npz", refrun.HARNESS_DEN2, refrun.HARNESS_GPU_DEN2)] + [
         # the quadratic, Gaussian, quintic-spline and Marquis kernels
         (f"edge3600_pc10_k{k}.npz", refrun.HARNESS_SKF[k], refrun.HARNESS_GPU_SKF[k])
         for k in (1, 2, 3, 7, 8)] + [
         # smoothing length kept / Benz's rate (the reference's own
         # smoothing_length and mass_evolution on the host, the pairs and
         # gradients from the shim), XSPH and the artificial viscosities
         # inside the GPU single_step, the Wang and Marquis boundary forces
         (_input_of(t), refrun.HARNESS_VARIANT[t], refrun.HARNESS_GPU_VARIANT[t])
         for t in refrun.VARIANT_TAGS]


@pytest.mark.parametrize("name,cpu_exe,gpu_exe", CASES)
def test_reference_time_integration_with_gpu_single_step(name, cpu_exe, gpu_exe):
    if not (os.path.exists(cpu_exe) and os.path.exists(gpu_exe)):
        pytest.skip("oracle/_ref harnesses not built (oracle/build_ref.sh)")
    print(f"[gpu] drop-in harness {name}", flush=True)
    g = load(name)
    sc = scenario(g)
    t_end = float(g["integration_time"])
    cpu = refrun.run(refrun.MODE_TIME_INTEGRATION, sc, integration_time=t_end,
                     harness=cpu_exe)
    gpu = refrun.run(refrun.MODE_TIME_INTEGRATION, sc, integration_time=t_end,
                     harness=gpu_exe)
    assert gpu["itimestep"] == cpu["itimestep"] == int(g["out_itimestep"])
    assert gpu["nip"] == cpu["nip"] and gpu["bnip"] == cpu["bnip"]
    for k in ("x", "vx", "rho", "hsml", "A", "h", "eps", "sigma", "p", "zeta"):
        a, b = gpu[k], cpu[k]
        err = float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
        assert err <= 1e-12, (k, err)
