"""The Hibler and Gutfraind pressures (eos_alg = 1, 2; eos.f90:59, 94-102)
on the device: bit for bit against the reference built with each PARAMETER
(golden vectors of tests/golden/make_golden.py), within the mixed-precision
tolerance against the oracle, and per instance in a packed ensemble."""
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


@pytest.mark.parametrize("eos", [1, 2])
def test_pressure_switch_single_step_fp64_bit_exact_vs_reference_golden(eos):
    say(f"single_step fp64 eos_alg={eos} golden box2500")
    g = load(f"box2500_single_step_eos{eos}.npz")
    sc = scenario(g)
    assert sc.eos_alg == eos
    out = Simp.for_scenario(sc).single_step(sc)
    assert out["nip"] == int(g["out_nip"]) and out["bnip"] == int(g["out_bnip"])
    for k in ("dxdt", "dvxdt", "dhdt", "dAdt", "p", "sigma"):
        assert np.array_equal(out[k], g["out_" + k]), (k, nrel(out[k], g["out_" + k]))


@pytest.mark.parametrize("eos", [1, 2])
def test_pressure_switch_time_integration_fp64_vs_reference_golden(eos):
    say(f"time_integration fp64 eos_alg={eos} golden edge3600")
    g = load(f"edge3600_pc10_eos{eos}.npz")
    sc = scenario(g)
    assert sc.eos_alg == eos
    K = int(g["out_itimestep"])
    out = Simp.for_scenario(sc).time_integration(sc, K)
    assert out["itimestep"] == K
    for k in ("x", "vx", "rho", "hsml", "A", "h"):
        e = nrel(getattr(sc, k), g["out_" + k])
        assert e <= 1e-12, (k, e)


@pytest.mark.parametrize("eos", [1, 2])
def test_pressure_switch_mixed_single_step_within_tolerance(eos):
    say(f"single_step mixed eos_alg={eos} vortex 40k")
    sc = S.vortex()
    rng = np.random.default_rng(9)
    sc.vx = rng.normal(0.0, 0.05, size=(sc.n, 2))
    sc.h = sc.h * rng.uniform(0.97, 1.03, size=sc.n)
    sc.rho = sc.h * 900.0
    sc.time = 10.0
    sc.eos_alg = eos
    got = Simp.for_scenario(sc, precision=_lib.SIMP_MIXED).single_step(sc)
    want = capi.Oracle().single_step(sc)
    for k in ("dvxdt", "dhdt", "dAdt", "p", "sigma"):
        e = nrel(got[k], want[k])
        assert e <= 1e-4, (k, e)


@pytest.mark.parametrize("precision", [_lib.SIMP_FP64, _lib.SIMP_MIXED])
def test_pressure_switches_per_instance_in_an_ensemble(precision):
    say(f"ensemble with eos_alg 3, 1, 2 members, precision={precision}")
    scs = S.ensemble(count=3, nmin=1500, nmax=4000, seed=53)
    for sc, eos in zip(scs, (3, 1, 2)):
        sc.time = 1e-6
        sc.eos_alg = eos
    ens = Ensemble(precision)
    ids = ens.add(scs)
    ens.step(4)
    for sc, i in zip(scs, ids):
        got = ens.download(i)
        # each member alone on a handle of its own: the same bits
        a = sc.copy()
        Simp.for_scenario(a, precision=precision).time_integration(a, 4)
        for k in ("x", "vx", "rho", "hsml", "h", "A"):
            assert np.array_equal(got[k], getattr(a, k)), (sc.eos_alg, k)
        b = sc.copy()
        capi.Oracle().run_pc(b, 4)
        tol = 1e-12 if precision == _lib.SIMP_FP64 else 1e-4
        for k in ("x", "vx", "rho", "hsml", "h", "A"):
            assert nrel(got[k], getattr(b, k)) <= tol, (sc.eos_alg, k)
        # the displacement against the oracle's, not only the absolute
        # positions (which a wrong step would barely move); a few ulps of x
        # allowed on top (device exp may differ in the last bit)
        d_got = got["x"] - sc.x
        d_want = b.x - sc.x
        rel = 1e-9 if precision == _lib.SIMP_FP64 else 1e-4
        bound = rel * np.max(np.abs(d_want)) + 8 * np.spacing(np.max(np.abs(sc.x)))
        assert np.max(np.abs(d_got - d_want)) <= bound, sc.eos_alg
