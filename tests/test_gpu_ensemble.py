"""Packed ensembles: every instance's result is the same alone and in the
mix (bit for bit: the kernels never combine instances), step budgets, adding
and removing instances between steps, and failure isolation."""
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

FIELDS = ("x", "vx", "rho", "hsml", "h", "A")


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _members(count=4, seed=41):
    scs = S.ensemble(count=count, nmin=1500, nmax=7000, seed=seed)
    for sc in scs:
        sc.time = 1e-6          # forcing on from the first step
    return scs


def _alone(sc, steps, precision):
    a = sc.copy()
    Simp.for_scenario(a, precision=precision).time_integration(a, steps)
    return a


@pytest.mark.parametrize("precision", [_lib.SIMP_FP64, _lib.SIMP_MIXED])
def test_each_instance_is_the_same_alone_and_in_the_mix(precision):
    say(f"ensemble alone-vs-mix precision={precision}")
    scs = _members()
    ens = Ensemble(precision)
    ids = ens.add(scs)
    ens.step(6)
    for sc, i in zip(scs, ids):
        got = ens.download(i)
        want = _alone(sc, 6, precision)
        for k in FIELDS:
            assert np.array_equal(got[k], getattr(want, k)), (i, k)
        assert got["time"] == want.time and got["itimestep"] == 6


def test_ensemble_member_matches_the_oracle():
    say("ensemble member vs oracle")
    scs = _members(count=3, seed=43)
    ens = Ensemble(_lib.SIMP_FP64)
    ids = ens.add(scs)
    ens.step(5)
    for sc, i in zip(scs, ids):
        got = ens.download(i)
        b = sc.copy()
        capi.Oracle().run_pc(b, 5)
        for k in FIELDS:
            want = getattr(b, k)
            err = np.max(np.abs(got[k] - want)) / max(np.max(np.abs(want)), 1e-300)
            assert err <= 1e-12, (i, k, err)


def test_step_budgets_and_adding_and_removing_between_steps():
    say("ensemble budgets / add / remove")
    a, b, c = _members(count=3, seed=47)
    prec = _lib.SIMP_FP64
    ens = Ensemble(prec)
    ia, ib = ens.add([a, b], nsteps=[3, 0])
    ens.step(5)
    st = ens.status(ia)
    assert st["itimestep"] == 3 and not st["active"] and st["status"] == _lib.SIMP_OK
    assert ens.status(ib)["itimestep"] == 5
    (ic,) = ens.add([c], nsteps=[0])
    ens.step(4)
    got_a = ens.download(ia)
    ens.remove([ia])
    ens.step(2)
    assert ens.count()[0] == 2
    for sc, i, k in ((b, ib, 11), (c, ic, 6)):
        got = ens.download(i)
        want = _alone(sc, k, prec)
        for f in FIELDS:
            assert np.array_equal(got[f], getattr(want, f)), (i, f)
    want_a = _alone(a, 3, prec)
    for f in FIELDS:
        assert np.array_equal(got_a[f], getattr(want_a, f)), ("a", f)


def test_a_failing_instance_does_not_disturb_the_others():
    say("ensemble failure isolation")
    scs = _members(count=3, seed=53)
    bad = scs[1].copy()
    bad.mass[7] = np.nan            # non-finite state: flagged by the device checks
    ens = Ensemble(_lib.SIMP_FP64)
    ids = ens.add([scs[0], bad, scs[2]])
    ens.step(4)
    st = ens.status(ids[1])
    assert st["status"] == _lib.SIMP_E_NAN and not st["active"]
    for sc, i in ((scs[0], ids[0]), (scs[2], ids[2])):
        assert ens.status(i)["status"] == _lib.SIMP_OK
        got = ens.download(i)
        want = _alone(sc, 4, _lib.SIMP_FP64)
        for f in FIELDS:
            assert np.array_equal(got[f], getattr(want, f)), (i, f)


def test_summation_density_members_across_a_repack():
    say("ensemble den_alg=1 members across add")
    scs = _members(count=3, seed=59)
    for sc in scs:
        sc.den_alg = 1
    prec = _lib.SIMP_FP64
    ens = Ensemble(prec)
    ia, ib = ens.add(scs[:2])
    ens.step(3)
    (ic,) = ens.add([scs[2]])      # the repack carries kernel_zero of a and b
    ens.step(3)
    for sc, i, k in ((scs[0], ia, 6), (scs[1], ib, 6), (scs[2], ic, 3)):
        got = ens.download(i)
        want = _alone(sc, k, prec)
        for f in FIELDS:
            assert np.array_equal(got[f], getattr(want, f)), (i, f)
