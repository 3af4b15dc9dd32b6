"""Failure paths the reference reports with WRITE + STOP, forced from the
host with legitimate inputs (every kernel still runs to completion):

* more than max_ind_int = 200 interactions of one particle at a step that
  is a multiple of print_step = 2000 (``nnps.f90:512-515,660-668``):
  SIMP_E_CAPACITY, and not at other steps;
* the same in a packed batch: only that instance stops, the others go on;
* a served request whose footprint can never fit the HBM budget is refused
  with SIMP_E_CAPACITY while the others complete.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from oracle import capi
from simp_amd import _lib
from simp_amd import scenarios as S
from simp_amd.api import Simp
from simp_amd.ensemble import Ensemble, footprint
from simp_amd.serve import Request, Server

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _dense(hsml_factor):
    # 30 x 30 particles 100 m apart with hsml 9 x spacing: up to ~260
    # interactions per particle (7 x spacing: at most ~160)
    return S.box(nside=30, spacing=100.0, seed=3, hsml_factor=hsml_factor)


def _max_interactions(sc):
    P = capi.Oracle().direct_find(sc.x, sc.hsml, sc.xb)
    cnt = np.bincount(P.p_i - 1, minlength=sc.n) + np.bincount(P.p_j - 1, minlength=sc.n)
    return int((cnt + np.bincount(P.bp_j - 1, minlength=sc.n)).max())


def test_more_than_200_interactions_stop_at_the_check_step():
    sc = _dense(9.0)
    assert _max_interactions(sc) > 200
    say("capacity: single_step at itimestep 1999 and 2000")
    dev = Simp.for_scenario(sc)
    dev.single_step(sc, itimestep=1999)          # not a check step
    with pytest.raises(_lib.SimpError) as ei:
        dev.single_step(sc, itimestep=2000)
    assert ei.value.status == _lib.SIMP_E_CAPACITY
    ok = _dense(7.0)
    assert _max_interactions(ok) <= 200
    Simp.for_scenario(ok).single_step(ok, itimestep=2000)


def test_capacity_stop_during_time_integration():
    sc = _dense(9.0)
    say("capacity: device-resident steps across itimestep 2000")
    dev = Simp.for_scenario(sc)
    dev.upload(sc, itimestep=1996)
    dev.step(3)                                    # steps 1996 .. 1998: fine
    with pytest.raises(_lib.SimpError) as ei:
        dev.step(3)
    assert ei.value.status == _lib.SIMP_E_CAPACITY


def test_capacity_stop_isolated_in_a_batch():
    bad = _dense(9.0)
    good = [S.box_member(3000, 80e3, (6.0, -2.0), seed=300 + k) for k in range(2)]
    say("capacity: one instance of three in a batch, steps 1998 .. 2001")
    ens = Ensemble(_lib.SIMP_FP64)
    ids = ens.add([good[0], bad, good[1]], itimestep=1998)
    ens.step(4)
    st = [ens.status(i) for i in ids]
    assert st[1]["status"] == _lib.SIMP_E_CAPACITY and not st[1]["active"]
    for k in (0, 2):
        assert st[k]["status"] == _lib.SIMP_OK and st[k]["active"]
        assert st[k]["itimestep"] == 2002
    ens.close()


def test_served_request_that_can_never_fit():
    small = [S.box_member(3000, 80e3, (5.0, 1.0), seed=400 + k) for k in range(3)]
    big = S.box_member(60000, 200e3, (5.0, 1.0), seed=499)
    budget = 2 * max(footprint(sc) for sc in small)
    assert footprint(big) > budget
    reqs = [Request(0, 0.0, small[0], 20), Request(1, 0.0, big, 20),
            Request(2, 0.0, small[1], 15), Request(3, 0.0, small[2], 10)]
    say("serve: a request larger than the budget")
    m = Server(precision=_lib.SIMP_MIXED, budget_bytes=budget, chunk=5).run(reqs)
    res = {r.rid: r for r in m["results"]}
    assert res[1].status == _lib.SIMP_E_CAPACITY
    assert m["completed"] == 3 and m["failed"] == 1


def _oracle_stats(sc):
    P = capi.Oracle().nnps(sc.x, sc.hsml, sc.xb)
    cnt = (np.bincount(P.p_i - 1, minlength=sc.n) + np.bincount(P.p_j - 1, minlength=sc.n)
           + np.bincount(P.bp_j - 1, minlength=sc.n))
    avg = float(np.float32(cnt.sum()) / np.float32(sc.n))
    return {"max": int(cnt.max()), "min": int(cnt.min()), "average": avg,
            "no_interaction": int((cnt == 0).sum()), "nip": int(len(P.p_i))}


@pytest.mark.parametrize("case", ["box2500", "edge3600", "dense"])
def test_interaction_stats_match_the_reference_lists(case):
    """interaction_stats (nnps.f90:613-658) of a single_step: max, min and
    average interactions per particle, particles with none, pairs -- from
    the oracle's pair lists of the same state."""
    if case == "box2500":
        sc = S.box()
    elif case == "edge3600":
        sc = S.ice_edge(nside=60, spacing=1000.0, seed=3)
        sc.hsml[::97] *= 0.3          # a few particles without any neighbour
        sc.max_hsml = float(sc.hsml.max() * 10)
    else:
        sc = _dense(7.0)
    say(f"interaction stats {case}")
    dev = Simp.for_scenario(sc)
    with pytest.raises(_lib.SimpError) as ei:
        dev.interaction_stats()
    assert ei.value.status == _lib.SIMP_E_STATE
    dev.single_step(sc, itimestep=7)
    got = dev.interaction_stats()
    want = _oracle_stats(sc)
    assert got["itimestep"] == 7
    for k in want:
        assert got[k] == want[k], (k, got[k], want[k])
    if case == "edge3600":
        assert want["no_interaction"] > 0


def test_interaction_stats_collected_at_print_steps():
    sc = S.box_member(3000, 80e3, (6.0, -2.0), seed=310)
    say("interaction stats: steps across itimestep 2000 in a batch")
    ens = Ensemble(_lib.SIMP_FP64)
    (i,) = ens.add([sc], itimestep=1998)
    ens.step(1)
    with pytest.raises(_lib.SimpError):
        ens.interaction_stats(i)           # no print step evaluated yet
    ens.step(3)                             # 1999, 2000, 2001
    st = ens.interaction_stats(i)
    assert st["itimestep"] == 2000 and st["max"] > 0 and st["nip"] > sc.n
    ens.close()
