"""The stress_2 particle approximation (pa_sph = 2,
internal_stress.f90:189-216) on the device: bit for bit against the
reference built with that PARAMETER (golden vectors of
tests/golden/make_golden.py), within the mixed-precision tolerance against
the oracle, and per instance in a packed ensemble next to stress_1
members."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from golden_io import load, scenario
from oracle import capi
from simp_amd import _lib
from simp_amd import scenarios as S
from simp_amd.api import Simp
from simp_amd.ensemble import Ensemble

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def nrel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_stress_2_single_step_fp64_bit_exact_vs_reference_golden():
    say("single_step fp64 stress_2 golden box2500")
    g = load("box2500_single_step_pa2.npz")
    sc = scenario(g)
    assert sc.pa_sph == 2
    out = Simp.for_scenario(sc).single_step(sc)
    assert out["nip"] == int(g["out_nip"]) and out["bnip"] == int(g["out_bnip"])
    for k in ("dxdt", "dvxdt", "dhdt", "dAdt", "sigma"):
        assert np.array_equal(out[k], g["out_" + k]), (k, nrel(out[k], g["out_" + k]))


def test_stress_2_time_integration_fp64_vs_reference_golden():
    say("time_integration fp64 stress_2 golden edge3600")
    g = load("edge3600_pc10_pa2.npz")
    sc = scenario(g)
    K = int(g["out_itimestep"])
    dev = Simp.for_scenario(sc)
    out = dev.time_integration(sc, K)
    assert out["itimestep"] == K
    for k in ("x", "vx", "rho", "hsml", "A", "h"):
        e = nrel(getattr(sc, k), g["out_" + k])
        assert e <= 1e-12, (k, e)


def test_stress_2_mixed_single_step_within_tolerance():
    say("single_step mixed stress_2 vortex 40k")
    sc = S.vortex()
    sc.vx = np.random.default_rng(8).normal(0.0, 0.05, size=(sc.n, 2))
    sc.time = 10.0
    sc.pa_sph = 2
    got = Simp.for_scenario(sc, precision=_lib.SIMP_MIXED).single_step(sc)
    want = capi.Oracle().single_step(sc)
    for k in ("dvxdt", "dhdt", "dAdt", "sigma"):
        e = nrel(got[k], want[k])
        assert e <= 1e-4, (k, e)


@pytest.mark.parametrize("precision", [_lib.SIMP_FP64, _lib.SIMP_MIXED])
def test_stress_2_member_in_an_ensemble(precision):
    say(f"ensemble with stress_1 and stress_2 members, precision={precision}")
    scs = S.ensemble(count=3, nmin=1500, nmax=4000, seed=47)
    for i, sc in enumerate(scs):
        sc.time = 1e-6
        sc.pa_sph = 2 if i % 2 == 0 else 1
    ens = Ensemble(precision)
    ids = ens.add(scs)
    ens.step(4)
    for sc, i in zip(scs, ids):
        got = ens.download(i)
        a = sc.copy()
        Simp.for_scenario(a, precision=precision).time_integration(a, 4)
        for k in ("x", "vx", "rho", "hsml", "h", "A"):
            assert np.array_equal(got[k], getattr(a, k)), (i, sc.pa_sph, k)
    if precision == _lib.SIMP_FP64:
        b = scs[0].copy()
        capi.Oracle().run_pc(b, 4)
        got = ens.download(ids[0])
        for k in ("x", "vx", "rho", "hsml", "h", "A"):
            assert nrel(got[k], getattr(b, k)) <= 1e-12, k
