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


@pytest.mark.skipif(not (os.path.exists(refrun.HARNESS) and os.path.exists(refrun.HARNESS_GPU)),
                    reason="oracle/_ref harnesses not built (oracle/build_ref.sh)")
@pytest.mark.parametrize("name", ["box2500_pc20.npz", "edge3600_pc10.npz"])
def test_reference_time_integration_with_gpu_single_step(name):
    print(f"[gpu] drop-in harness {name}", flush=True)
    g = load(name)
    sc = scenario(g)
    t_end = float(g["integration_time"])
    cpu = refrun.run(refrun.MODE_TIME_INTEGRATION, sc, integration_time=t_end)
    gpu = refrun.run(refrun.MODE_TIME_INTEGRATION, sc, integration_time=t_end,
                     harness=refrun.HARNESS_GPU)
    assert gpu["itimestep"] == cpu["itimestep"] == int(g["out_itimestep"])
    assert gpu["nip"] == cpu["nip"] and gpu["bnip"] == cpu["bnip"]
    for k in ("x", "vx", "rho", "hsml", "A", "h", "eps", "sigma", "p", "zeta"):
        a, b = gpu[k], cpu[k]
        err = float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
        assert err <= 1e-12, (k, err)
