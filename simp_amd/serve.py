"""Continuous-batching request loop over a packed :class:`Ensemble`.

Requests (one simulation each: a scenario and a number of steps) arrive over
time.  The loop admits arrived requests in arrival order while their device
footprint fits the HBM budget, steps every admitted simulation together in
chunks of ``chunk`` steps (one packed launch per kernel), and retires each
simulation as soon as its own step budget is used up: its final state is
copied to the host, its memory is released, and waiting requests take its
place at the next chunk boundary.  Requests that arrive while others are on
the GPU therefore join the running batch instead of waiting for it to drain.

Reported: throughput, p50/p99 step latency (wall time of one step of the
running batch, as seen by every simulation in it), p50/p99 queueing delay and
completion latency (arrival to final state on the host).  The GPU-busy time
counts the step launches and the layout updates of admission and retirement
(one per step boundary that has any).
"""
from __future__ import annotations

import dataclasses
import math
import time
from collections import deque
from typing import Callable, List, Optional

import numpy as np

from . import _lib
from . import scenarios as S
from .ensemble import Ensemble, footprint


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float          # seconds after the start of the run
    scenario: S.Scenario
    steps: int


@dataclasses.dataclass
class Result:
    rid: int
    n: int
    steps: int
    arrival: float
    admitted: float = math.nan
    finished: float = math.nan
    status: int = _lib.SIMP_OK
    state: Optional[dict] = None


def poisson_requests(count: int = 400, rate: float = 20.0, nmin: int = 5000,
                     nmax: int = 250000, smin: int = 100, smax: int = 2000,
                     seed: int = 5) -> List[Request]:
    """Poisson arrivals (rate per second) of closed-box simulations with
    lognormal particle counts clipped to [nmin, nmax], step counts uniform in
    [smin, smax], box size U[50, 300] km and uniform wind U[5, 20] m/s of
    random direction."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(count):
        t += rng.exponential(1.0 / rate)
        N = float(np.clip(rng.lognormal(math.log(30000.0), 0.9), nmin, nmax))
        L = rng.uniform(50e3, 300e3)
        ang = rng.uniform(0, 2 * math.pi)
        spd = rng.uniform(5.0, 20.0)
        sc = S.box_member(N, L, (spd * math.cos(ang), spd * math.sin(ang)),
                          seed * 1000003 + k)
        out.append(Request(k, t, sc, int(rng.integers(smin, smax + 1))))
    return out


class Server:
    """``ensemble``: the packed batch to serve from (default: a new
    :class:`Ensemble` of the given precision); any object with the same
    ``update`` / ``step`` / ``list`` / ``download`` methods will do."""

    def __init__(self, precision: int = _lib.SIMP_MIXED, budget_bytes: int = 8 << 30,
                 chunk: int = 10, keep_states: bool = False,
                 clock: Callable[[], float] = time.perf_counter, ensemble=None):
        self.ens = ensemble if ensemble is not None else Ensemble(precision)
        self.budget = int(budget_bytes)
        self.chunk = int(chunk)
        self.keep = keep_states
        self.clock = clock

    def run(self, requests: List[Request]) -> dict:
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        results = {r.rid: Result(r.rid, r.scenario.n, r.steps, r.arrival) for r in requests}
        active = {}                  # ensemble id -> (request, footprint)
        retire = []                  # finished ids, removed with the next layout update
        used = 0
        step_lat = []
        busy = 0.0                   # time in step launches
        relayout = 0.0               # time in layout updates (admission / retirement)
        updates = 0
        work = 0
        t0 = self.clock()
        while pending or active:
            now = self.clock() - t0
            admit = []
            while pending and pending[0].arrival <= now:
                r = pending[0]
                fp = footprint(r.scenario)
                if fp > self.budget:                 # can never fit
                    results[r.rid].status = _lib.SIMP_E_CAPACITY
                    results[r.rid].finished = now
                    pending.popleft()
                    continue
                if used + fp > self.budget:           # wait for memory (FIFO)
                    break
                admit.append((r, fp))
                used += fp
                pending.popleft()
            if admit or (retire and (active or pending)):
                # one rebuild of the packed layout per step boundary: the
                # finished simulations leave and the admitted ones join
                ts = self.clock()
                ids = self.ens.update(retire, [r.scenario for r, _ in admit],
                                      [r.steps for r, _ in admit])
                relayout += self.clock() - ts
                updates += 1
                retire = []
                tadm = self.clock() - t0
                for i, (r, fp) in zip(ids, admit):
                    active[i] = (r, fp)
                    results[r.rid].admitted = tadm
            if not active:
                if pending:
                    time.sleep(max(0.0, pending[0].arrival - (self.clock() - t0)))
                continue
            ts = self.clock()
            self.ens.step(self.chunk)
            dt = self.clock() - ts
            busy += dt
            step_lat.append(dt / self.chunk)
            status = self.ens.list()
            done = [i for i, st in status.items() if not st["active"] and i in active]
            for i in done:
                st = status[i]
                r, fp = active.pop(i)
                res = results[r.rid]
                res.status = st["status"]
                if self.keep:
                    res.state = self.ens.download(i)
                res.finished = self.clock() - t0
                work += r.scenario.n * (st["itimestep"])
                used -= fp
                retire.append(i)
        wall = self.clock() - t0
        out = self._metrics(results, step_lat, busy + relayout, work, wall)
        out.update(step_s=busy, relayout_s=relayout, relayouts=updates,
                   step_throughput=work / busy if busy > 0 else 0.0)
        return out

    @staticmethod
    def _metrics(results, step_lat, busy, work, wall) -> dict:
        res = list(results.values())
        ok = [r for r in res if r.status == _lib.SIMP_OK and not math.isnan(r.finished)]
        comp = np.array([r.finished - r.arrival for r in ok]) if ok else np.zeros(1)
        wait = np.array([r.admitted - r.arrival for r in ok]) if ok else np.zeros(1)
        lat = np.array(step_lat) if step_lat else np.zeros(1)
        return {
            "requests": len(res), "completed": len(ok),
            "failed": sum(1 for r in res if r.status != _lib.SIMP_OK),
            "particle_steps": int(work), "wall_s": wall, "gpu_busy_s": busy,
            "throughput": work / wall if wall > 0 else 0.0,
            "busy_throughput": work / busy if busy > 0 else 0.0,
            "p50_step_ms": 1e3 * float(np.percentile(lat, 50)),
            "p99_step_ms": 1e3 * float(np.percentile(lat, 99)),
            "p50_wait_s": float(np.percentile(wait, 50)),
            "p99_wait_s": float(np.percentile(wait, 99)),
            "p50_completion_s": float(np.percentile(comp, 50)),
            "p99_completion_s": float(np.percentile(comp, 99)),
            "results": res,
        }
