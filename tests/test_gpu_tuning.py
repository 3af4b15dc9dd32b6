"""The tuning parameters and the autotuner (no counterpart in the
reference): no setting changes a result, so every setting must give the
same bits as the default, in both precisions; the autotuner keeps a setting
at least as fast as the one it started from, and its trial steps are steps
of the simulation, the same as steps with the default setting."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from simp_amd import _lib
from simp_amd import scenarios as S
from simp_amd.api import Simp
from simp_amd.ensemble import Ensemble

pytestmark = pytest.mark.gpu

FIELDS = ("x", "vx", "rho", "hsml", "h", "A")
SETTINGS = [dict(passA_batch=1), dict(passA_batch=3), dict(passA_batch=4),
            dict(passB_batch=2), dict(passB_batch=3), dict(passB_batch=4),
            dict(skin=0.005), dict(skin=0.04), dict(passA_batch=1, passB_batch=2, skin=0.2)]


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _moving_box():
    sc = S.box(seed=35)
    sc.vx = np.random.default_rng(12).normal(0.0, 5.0, size=sc.vx.shape)
    sc.tse = 0
    sc.dt = 0.05
    sc.time = 1.0
    return sc


def _run(sc, precision, steps, **tuning):
    dev = Simp.for_scenario(sc, precision=precision)
    if tuning:
        dev.set_tuning(**tuning)
    dev.upload(sc)
    dev.step(steps)
    out = dev.download()
    out["search"] = dev.search_stats()
    dev.close()
    return out


@pytest.mark.parametrize("precision", [_lib.SIMP_FP64, _lib.SIMP_MIXED])
def test_every_setting_gives_the_default_bits(precision):
    sc = _moving_box()
    say(f"tuning settings vs default, box2500 moving, 16 steps, precision={precision}")
    want = _run(sc, precision, 16)
    for t in SETTINGS:
        got = _run(sc, precision, 16, **t)
        for k in FIELDS:
            assert np.array_equal(got[k], want[k]), (t, k)
    # the skin changes how often the candidate lists are rebuilt
    small = _run(sc, precision, 16, skin=0.005)["search"]["candidate_rebuilds"]
    large = _run(sc, precision, 16, skin=0.2)["search"]["candidate_rebuilds"]
    assert small > large


def test_autotuner_keeps_a_faster_setting_and_steps_the_simulation():
    sc = S.ice_edge(nside=200, spacing=1000.0, seed=3)
    sc.time = 1e-6
    say(f"autotune ice edge {sc.n} particles, mixed")
    dev = Simp.for_scenario(sc, precision=_lib.SIMP_MIXED)
    dev.upload(sc)
    res = dev.tune(10)
    assert 1 <= res["passA_batch"] <= 4 and res["passB_batch"] in (0, 2, 3, 4)
    assert 0 < res["skin"] <= 0.2
    assert 0 < res["us_best"] <= res["us_start"]
    assert dev.get_tuning()["passA_batch"] == res["passA_batch"]
    dev.step(5)
    got = dev.download()
    dev.close()
    # the same number of steps with the default setting
    steps = got["itimestep"]
    assert steps >= 11 * 10
    want = _run(sc, _lib.SIMP_MIXED, steps)
    for k in FIELDS:
        assert np.array_equal(got[k], want[k]), k


def test_autotuner_on_a_packed_batch():
    scs = S.ensemble(count=6, nmin=2000, nmax=6000, seed=61)
    for sc in scs:
        sc.time = 1e-6
    say("autotune a packed batch of 6")
    ens = Ensemble(_lib.SIMP_FP64)
    ids = ens.add(scs)
    res = ens.tune(5)
    assert 0 < res["us_best"] <= res["us_start"]
    ens.step(3)
    for sc, i in zip(scs, ids):
        got = ens.download(i)
        want = _run(sc, _lib.SIMP_FP64, got["itimestep"])
        for k in FIELDS:
            assert np.array_equal(got[k], want[k]), k
    ens.close()
