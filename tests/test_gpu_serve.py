"""The continuous-batching request loop: requests that arrive while others
run join the batch, each finishes after exactly its own step count, and its
result is the same as running it alone."""
from __future__ import annotations

import numpy as np
import pytest

from simp_amd import _lib
from simp_amd import scenarios as S
from simp_amd.api import Simp
from simp_amd.serve import Request, Server

pytestmark = pytest.mark.gpu


class StepClock:
    """A clock that advances by a fixed amount per reading: arrivals are then
    a function of the loop's progress only (deterministic admission)."""

    def __init__(self, tick=0.01):
        self.t = 0.0
        self.tick = tick

    def __call__(self):
        self.t += self.tick
        return self.t


def test_served_results_equal_runs_alone():
    print("[gpu] serve loop alone-vs-mix", flush=True)
    reqs = []
    for k, (N, L, steps, arr) in enumerate([(2000, 60e3, 25, 0.0), (4500, 120e3, 40, 0.0),
                                            (3000, 80e3, 15, 0.05), (6000, 200e3, 30, 0.1),
                                            (2500, 50e3, 20, 0.15)]):
        sc = S.box_member(N, L, (6.0, -3.0 + k), seed=900 + k)
        sc.time = 1e-6
        reqs.append(Request(k, arr, sc, steps))
    srv = Server(precision=_lib.SIMP_MIXED, budget_bytes=1 << 30, chunk=5, keep_states=True,
                 clock=StepClock())
    m = srv.run(reqs)
    assert m["completed"] == len(reqs) and m["failed"] == 0
    res = {r.rid: r for r in m["results"]}
    # requests were admitted while earlier ones were still running
    joined = [(a.rid, b.rid) for a in res.values() for b in res.values()
              if a.admitted < b.admitted < a.finished]
    assert joined, [(r.rid, r.admitted, r.finished) for r in res.values()]
    for r in reqs:
        got = res[r.rid].state
        assert got["itimestep"] == r.steps
        alone = r.scenario.copy()
        Simp.for_scenario(alone, precision=_lib.SIMP_MIXED).time_integration(alone, r.steps)
        for f in ("x", "vx", "rho", "hsml", "h", "A"):
            assert np.array_equal(got[f], getattr(alone, f)), (r.rid, f)


def test_admission_waits_for_memory():
    print("[gpu] serve admission", flush=True)
    scs = [S.box_member(3000, 80e3, (5.0, 0.0), seed=950 + k) for k in range(3)]
    from simp_amd.ensemble import footprint
    fp = max(footprint(sc) for sc in scs)
    reqs = [Request(k, 0.0, sc, 10) for k, sc in enumerate(scs)]
    srv = Server(precision=_lib.SIMP_FP64, budget_bytes=int(1.5 * fp), chunk=5,
                 clock=StepClock())
    m = srv.run(reqs)
    assert m["completed"] == 3
    res = sorted(m["results"], key=lambda r: r.rid)
    # only one fits at a time: each starts after the previous one finished
    assert res[1].admitted >= res[0].finished and res[2].admitted >= res[1].finished
