"""The request loop's host logic, on the CPU, with a stand-in for the packed
batch: admission by HBM budget in arrival order, a request that can never fit
is refused with SIMP_E_CAPACITY without blocking the queue, each request
finishes after exactly its own step count, and each step boundary rebuilds
the layout at most once (retirements and admissions together)."""
from __future__ import annotations

from simp_amd import _lib
from simp_amd import scenarios as S
from simp_amd.ensemble import footprint
from simp_amd.serve import Request, Server


class FakeBatch:
    def __init__(self):
        self.inst = {}
        self.next = 0
        self.updates = []

    def update(self, remove, scenarios, nsteps):
        self.updates.append((sorted(remove), len(scenarios)))
        for i in remove:
            del self.inst[i]
        ids = []
        for sc, k in zip(scenarios, nsteps):
            self.inst[self.next] = {"left": k, "it": 0}
            ids.append(self.next)
            self.next += 1
        return ids

    def step(self, n):
        for v in self.inst.values():
            a = min(n, v["left"])
            v["it"] += a
            v["left"] -= a

    def list(self):
        return {i: {"itimestep": v["it"], "steps_left": v["left"], "active": v["left"] > 0,
                    "status": _lib.SIMP_OK} for i, v in self.inst.items()}

    def download(self, i):
        return {"itimestep": self.inst[i]["it"]}


class StepClock:
    def __init__(self, tick=0.01):
        self.t = 0.0
        self.tick = tick

    def __call__(self):
        self.t += self.tick
        return self.t


def _reqs():
    scs = [S.box_member(2000 + 500 * k, 60e3, (5.0, 0.0), seed=600 + k) for k in range(5)]
    steps = [20, 35, 10, 25, 15]
    arr = [0.0, 0.0, 0.05, 0.1, 0.12]
    return [Request(k, a, sc, s) for k, (sc, s, a) in enumerate(zip(scs, steps, arr))]


def test_never_fitting_request_is_refused_and_the_rest_complete():
    reqs = _reqs()
    fps = [footprint(r.scenario) for r in reqs]
    big = S.box_member(40000, 200e3, (5.0, 0.0), seed=699)
    budget = 2 * max(fps)
    assert footprint(big) > budget
    reqs.insert(1, Request(99, 0.0, big, 10))
    fb = FakeBatch()
    m = Server(budget_bytes=budget, chunk=5, keep_states=True, clock=StepClock(),
               ensemble=fb).run(reqs)
    res = {r.rid: r for r in m["results"]}
    assert res[99].status == _lib.SIMP_E_CAPACITY
    assert m["failed"] == 1 and m["completed"] == len(reqs) - 1
    for r in reqs:
        if r.rid != 99:
            assert res[r.rid].state["itimestep"] == r.steps
    assert m["gpu_busy_s"] == m["step_s"] + m["relayout_s"]


def test_one_layout_update_per_step_boundary():
    reqs = _reqs()
    fb = FakeBatch()
    fps = sorted(footprint(r.scenario) for r in reqs)
    budget = fps[-1] + fps[-2]       # any two fit, never three: a waiting request
    m = Server(budget_bytes=budget, chunk=5, clock=StepClock(), ensemble=fb).run(reqs)
    assert m["completed"] == len(reqs)
    assert m["relayouts"] == len(fb.updates)
    # joins as soon as one leaves: both in the same update
    assert any(rm and add for rm, add in fb.updates), fb.updates
    # every admitted request was eventually retired (but the last ones, which
    # leave with the batch)
    retired = [i for rm, _ in fb.updates for i in rm]
    assert len(set(retired)) == len(retired)
