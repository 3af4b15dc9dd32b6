"""The Monaghan time step (tse = 1, time_step.f90:39-125, without artificial
viscosity) on the device: against the oracle (itself bit for bit with the
reference built with tse = 1 over ten steps, test_oracle) with each time
integration scheme, in mixed precision, and per instance in a packed
ensemble next to members on the VP time step."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from oracle import capi
from simp_amd import _lib
from simp_amd import scenarios as S
from simp_amd.api import Simp
from simp_amd.ensemble import Ensemble

pytestmark = pytest.mark.gpu

FIELDS = ("x", "vx", "rho", "hsml", "A", "h")


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def nrel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _box(seed=43):
    sc = S.box(seed=seed)
    sc.vx = np.random.default_rng(seed).normal(0.0, 0.05, size=sc.vx.shape)
    sc.time = 1.0
    sc.tse = 1
    return sc


@pytest.mark.parametrize("scheme", [1, 2, 3])
@pytest.mark.parametrize("den_alg", [3, 1])
def test_monaghan_time_step_each_scheme_fp64_vs_oracle(scheme, den_alg):
    say(f"time_integration fp64 tse=1 int_algorithm={scheme} den_alg={den_alg} box2500, 6 steps")
    sc = _box()
    sc.int_algorithm = scheme
    sc.den_alg = den_alg
    a, b = sc.copy(), sc.copy()
    Simp.for_scenario(a).time_integration(a, 6)
    capi.Oracle().run_pc(b, 6)
    # the step's dt follows the state: every step after the first uses it
    assert b.dt != sc.dt
    assert abs(a.dt - b.dt) <= 1e-12 * b.dt and abs(a.time - b.time) <= 1e-12 * b.time
    for k in FIELDS:
        e = nrel(getattr(a, k), getattr(b, k))
        assert e <= 1e-12, (k, e)


def test_monaghan_time_step_mixed_precision_within_tolerance():
    say("time_integration mixed tse=1 ice edge 3600, 3 steps")
    sc = S.ice_edge(nside=60, spacing=1000.0, seed=3)
    sc.tse = 1
    a, b = sc.copy(), sc.copy()
    Simp.for_scenario(a, precision=_lib.SIMP_MIXED).time_integration(a, 3)
    capi.Oracle().run_pc(b, 3)
    assert abs(a.dt - b.dt) <= 1e-4 * b.dt
    for k in FIELDS:
        e = nrel(getattr(a, k), getattr(b, k))
        assert e <= 1e-4, (k, e)
    assert nrel(a.x - sc.x, b.x - sc.x) <= 1e-4


@pytest.mark.parametrize("precision", [_lib.SIMP_FP64, _lib.SIMP_MIXED])
def test_monaghan_and_vp_time_steps_per_instance_in_an_ensemble(precision):
    say(f"ensemble with tse 2, 1, 1 members, precision={precision}")
    scs = S.ensemble(count=3, nmin=1500, nmax=4000, seed=61)
    for sc, tse in zip(scs, (2, 1, 1)):
        sc.time = 1e-6
        sc.tse = tse
    scs[2].den_alg = 1
    ens = Ensemble(precision)
    ids = ens.add(scs)
    ens.step(5)
    for sc, i in zip(scs, ids):
        got = ens.download(i)
        a = sc.copy()
        Simp.for_scenario(a, precision=precision).time_integration(a, 5)
        for k in FIELDS:
            assert np.array_equal(got[k], getattr(a, k)), (sc.tse, k)
        assert got["dt"] == a.dt
        if precision == _lib.SIMP_FP64:
            b = sc.copy()
            capi.Oracle().run_pc(b, 5)
            assert abs(got["dt"] - b.dt) <= 1e-12 * b.dt
            for k in FIELDS:
                assert nrel(got[k], getattr(b, k)) <= 1e-12, (sc.tse, k)
