"""The batched and served workloads at their own sizes.

* Ensemble: 64 closed boxes, particle counts lognormal in [2k, 60k] (about
  1.06 M particles), box 50-300 km, uniform wind of random direction and
  speed U[5, 20] m/s, 500 steps in one packed launch per kernel.  Every
  member is the same alone and in the mix, bit for bit, in both precisions;
  the smallest, median and largest members match the FP64 oracle over a few
  steps; the mixed members stay within the mixed path's tolerance of FP64.
* Served: ten requests arriving while others run, N lognormal in
  [5k, 250k] with one request of at least 200k particles, 100-2,000 steps
  each, HBM-budget admission, mixed precision.  Every result is the same as
  running the request alone.
"""
from __future__ import annotations

import math
import sys

import numpy as np
import pytest

from oracle import capi
from simp_amd import _lib
from simp_amd import accuracy
from simp_amd import scenarios as S
from simp_amd.api import Simp
from simp_amd.ensemble import Ensemble
from simp_amd.serve import Request, Server

pytestmark = pytest.mark.gpu

FIELDS = ("x", "vx", "rho", "hsml", "h", "A")
STEPS = 500


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _alone(sc, steps, precision, diag=False):
    dev = Simp.for_scenario(sc, precision=precision)
    dev.set_diagnostics(diag)
    dev.upload(sc)
    dev.step(steps)
    out = dev.download()
    if diag:
        out["sigma"] = dev.download_diagnostics()["sigma"]
    dev.close()
    return out


@pytest.fixture(scope="module")
def members():
    scs = S.ensemble(count=64)
    assert min(sc.n for sc in scs) >= 2000 and max(sc.n for sc in scs) <= 60100
    return scs


@pytest.mark.parametrize("precision", [_lib.SIMP_MIXED, _lib.SIMP_FP64])
def test_ensemble_64_members_alone_equals_mix(members, precision):
    say(f"ensemble 64 members, {sum(s.n for s in members)} particles, {STEPS} steps, "
        f"precision={precision}")
    ens = Ensemble(precision)
    ids = ens.add(members)
    ens.step(STEPS)
    mixed_states = {}
    for k, (sc, i) in enumerate(zip(members, ids)):
        got = ens.download(i)
        assert got["itimestep"] == STEPS
        want = _alone(sc, STEPS, precision)
        for f in FIELDS:
            assert np.array_equal(got[f], want[f]), (k, sc.n, f)
        assert got["time"] == want["time"]
        mixed_states[k] = got
    ens.close()
    if precision == _lib.SIMP_MIXED:
        # the mixed members against FP64 runs of the same members
        order = np.argsort([sc.n for sc in members])
        for k in (order[0], order[len(order) // 2], order[-1]):
            sc = members[k]
            m = _alone(sc, STEPS, _lib.SIMP_MIXED, diag=True)
            f = _alone(sc, STEPS, _lib.SIMP_FP64, diag=True)
            tab = accuracy.field_errors(m, f, sc)
            say(f"  member {k} (n={sc.n}): worst norm {accuracy.worst(tab):.2e}")
            assert accuracy.worst(tab) <= 1e-4, (k, tab)


def test_ensemble_members_match_the_oracle(members):
    order = np.argsort([sc.n for sc in members])
    picks = [order[0], order[len(order) // 2], order[-1]]
    say(f"ensemble members {picks} vs oracle, 5 steps (packed with all 64)")
    ens = Ensemble(_lib.SIMP_FP64)
    ids = ens.add(members)
    ens.step(5)
    for k in picks:
        got = ens.download(ids[k])
        b = members[k].copy()
        capi.Oracle().run_pc(b, 5)
        for f in FIELDS:
            want = getattr(b, f)
            err = np.max(np.abs(got[f] - want)) / max(np.max(np.abs(want)), 1e-300)
            assert err <= 1e-12, (k, f, err)
    ens.close()


class StepClock:
    """Advances by a fixed amount per reading: arrivals then depend on the
    loop's progress only (deterministic admission)."""

    def __init__(self, tick=0.01):
        self.t = 0.0
        self.tick = tick

    def __call__(self):
        self.t += self.tick
        return self.t


def _requests():
    rng = np.random.default_rng(77)
    reqs, t = [], 0.0
    for k in range(10):
        t += rng.exponential(1.0 / 20.0)
        N = float(np.clip(rng.lognormal(math.log(30000.0), 0.9), 5000, 250000))
        if k == 3:
            N = 240000.0                  # one request near the top of the range
        L = rng.uniform(50e3, 300e3)
        ang = rng.uniform(0, 2 * math.pi)
        spd = rng.uniform(5.0, 20.0)
        sc = S.box_member(N, L, (spd * math.cos(ang), spd * math.sin(ang)), 7000 + k)
        reqs.append(Request(k, t, sc, int(rng.integers(100, 2001))))
    return reqs


def test_served_ten_requests_alone_equals_mix():
    reqs = _requests()
    ns = [r.scenario.n for r in reqs]
    assert max(ns) >= 200000 and min(ns) >= 5000
    say(f"serve 10 requests: n {min(ns)}..{max(ns)}, steps "
        f"{min(r.steps for r in reqs)}..{max(r.steps for r in reqs)}")
    srv = Server(precision=_lib.SIMP_MIXED, budget_bytes=16 << 30, chunk=10, keep_states=True,
                 clock=StepClock(0.02))
    m = srv.run(reqs)
    assert m["completed"] == len(reqs) and m["failed"] == 0
    res = {r.rid: r for r in m["results"]}
    joined = [(a.rid, b.rid) for a in res.values() for b in res.values()
              if a.admitted < b.admitted < a.finished]
    assert joined
    for r in reqs:
        got = res[r.rid].state
        assert got["itimestep"] == r.steps
        want = _alone(r.scenario, r.steps, _lib.SIMP_MIXED)
        for f in FIELDS:
            assert np.array_equal(got[f], want[f]), (r.rid, r.scenario.n, f)
