"""The reference's default scenario setup (in_out/initialize.f90
particle_placing, in_out/boundaries.f90 boundary_placing_1, the VP time
step): simp_amd.scenarios.from_maps reproduces, bit for bit and in order,
the initial state the compiled reference's own initialize produces from
its domain and boundary maps (golden vectors from
tests/golden/make_init_golden.py); on the device the scenario then steps
like the oracle."""
from __future__ import annotations

import numpy as np
import pytest

from golden_io import reference_maps
from simp_amd import scenarios as S


def test_from_maps_reproduces_the_reference_initialize():
    mp, g = reference_maps()
    sc = S.from_maps(mp)
    assert sc.n == 1696 and sc.nb == 600
    for k in ("x", "vx", "mass", "rho", "hsml", "A", "h", "xb", "massb"):
        assert np.array_equal(getattr(sc, k), g[k]), k
    assert sc.dt == float(g["dt"])
    assert sc.max_hsml == float(g["max_hsml"])


def test_coarser_resolution_and_bad_inputs():
    mp, _ = reference_maps()
    mp.resolution = 15000.0          # two particles per map cell and direction
    sc = S.from_maps(mp)
    assert sc.n == int(np.sum(mp.map == 1)) * 4
    mp.resolution = 1e9
    with pytest.raises(ValueError):
        S.from_maps(mp)


@pytest.mark.gpu
def test_reference_default_scenario_steps_like_the_oracle():
    from oracle import capi
    from simp_amd.api import Simp
    print("[gpu] reference default scenario, 20 steps fp64", flush=True)
    mp, _ = reference_maps()
    a = S.from_maps(mp)
    a.time = 1e-6                    # wind on (the forcing starts after t = 0)
    b = a.copy()
    Simp.for_scenario(a).time_integration(a, 20)
    capi.Oracle().run_pc(b, 20)
    for k in ("x", "vx", "rho", "hsml", "A", "h"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
