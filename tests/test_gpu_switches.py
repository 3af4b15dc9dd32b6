"""Mohr-Coulomb rheology (rheology = 2), the Wendland C2 / C4 kernels and
their boundary kernels (skf = 4, 5) and the Euler / leap-frog schemes
(int_algorithm = 1, 3) on the device, through the C ABI: bit for bit
against the reference built with each PARAMETER for one evaluation
(kernel values included), within 1e-12 over ten steps (the device's exp may
differ from the host's in the last bit), within the mixed path's tolerance
in mixed precision, combined with summation density against the oracle, and
per instance in a packed ensemble (alone equals mixed, bit for bit)."""
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

FIELDS = ("x", "vx", "rho", "hsml", "A", "h")


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def nrel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


# kernels whose value uses exp, tanh, cosh or log: the device's libm may
# differ from the host's in the last bit, so they are compared within 1e-12
TRANSCENDENTAL = ("k2", "k7", "k8")


@pytest.mark.parametrize("tag", ["mc", "c2", "c4", "k1", "k2", "k3", "k7", "k8"])
def test_switch_single_step_fp64_bit_exact_vs_reference_golden(tag):
    say(f"single_step fp64 {tag} golden box2500")
    g = load(f"box2500_single_step_{tag}.npz")
    sc = scenario(g)
    dev = Simp.for_scenario(sc)
    out = dev.single_step(sc)
    assert out["nip"] == int(g["out_nip"]) and out["bnip"] == int(g["out_bnip"])
    kv = dev.kernel_values(sc.n)
    for k in ("dxdt", "dvxdt", "dhdt", "dAdt", "eps", "zeta", "eta", "p", "sigma"):
        if tag in TRANSCENDENTAL:
            assert nrel(out[k], g["out_" + k]) <= 1e-12, (k, nrel(out[k], g["out_" + k]))
        else:
            assert np.array_equal(out[k], g["out_" + k]), (k, nrel(out[k], g["out_" + k]))
    if tag == "k8":
        # Marquis2's value cancels: for q in (2, 2.5) its terms of about 1
        # sum to about 1e-3, so a last-bit difference of log(cosh(.)) is
        # about 1e-6 of the value (the gradient, which the fields use, does
        # not cancel)
        w, want = kv["w"], g["out_w"]
        assert np.all(np.abs(w - want) <= 5e-6 * np.abs(want)), nrel(w, want)
    elif tag in TRANSCENDENTAL:
        assert nrel(kv["w"], g["out_w"]) <= 1e-12
    else:
        assert np.array_equal(kv["w"], g["out_w"])


@pytest.mark.parametrize("tag", ["mc", "c2", "c4", "k1", "k2", "k3", "k7", "k8", "euler", "leap",
                                 "tse1", "den2"])
def test_switch_time_integration_fp64_vs_reference_golden(tag):
    say(f"time_integration fp64 {tag} golden edge3600")
    g = load(f"edge3600_pc10_{tag}.npz")
    sc = scenario(g)
    K = int(g["out_itimestep"])
    dev = Simp.for_scenario(sc)
    dev.set_diagnostics(True)
    out = dev.time_integration(sc, K)
    assert out["itimestep"] == K
    assert abs(out["time"] - float(g["out_time"])) <= 1e-12 * abs(float(g["out_time"]))
    assert abs(out["dt"] - float(g["out_dt"])) <= 1e-12 * abs(float(g["out_dt"]))
    for k in FIELDS:
        e = nrel(getattr(sc, k), g["out_" + k])
        assert e <= 1e-12, (k, e)
    diag = dev.download_diagnostics()
    for k in ("eps", "sigma", "p", "zeta"):
        if np.any(g["out_" + k]):
            assert nrel(diag[k], g["out_" + k]) <= 1e-10, k
        else:
            assert not np.any(diag[k]), k


@pytest.mark.parametrize("scheme", [1, 3])
@pytest.mark.parametrize("den_alg", [3, 1, 2])
def test_schemes_with_both_densities_vs_oracle(scheme, den_alg):
    say(f"time_integration fp64 int_algorithm={scheme} den_alg={den_alg} box2500, 12 steps")
    sc = S.box(seed=41)
    sc.vx = np.random.default_rng(2).normal(0.0, 0.05, size=sc.vx.shape)
    sc.time = 1.0
    sc.int_algorithm = scheme
    sc.den_alg = den_alg
    a, b = sc.copy(), sc.copy()
    Simp.for_scenario(a).time_integration(a, 12)
    capi.Oracle().run_pc(b, 12)
    for k in FIELDS:
        e = nrel(getattr(a, k), getattr(b, k))
        assert e <= 1e-12, (k, e)


@pytest.mark.parametrize("sw", [{"rheology": 2}, {"skf": 4}, {"skf": 5}, {"skf": 1},
                                {"skf": 3}, {"int_algorithm": 1}, {"int_algorithm": 3},
                                {"den_alg": 2}, {"den_alg": 2, "int_algorithm": 3}])
def test_switch_mixed_precision_within_tolerance(sw):
    say(f"time_integration mixed {sw} vortex 40k, 20 steps")
    sc = S.vortex()
    sc.vx = np.random.default_rng(8).normal(0.0, 0.05, size=(sc.n, 2))
    sc.time = 10.0
    for k, v in sw.items():
        setattr(sc, k, v)
    a, b = sc.copy(), sc.copy()
    Simp.for_scenario(a, precision=_lib.SIMP_MIXED).time_integration(a, 20)
    capi.Oracle().run_pc(b, 20)
    for k in FIELDS:
        e = nrel(getattr(a, k), getattr(b, k))
        assert e <= 1e-4, (k, e)
    assert nrel(a.vx - sc.vx, b.vx - sc.vx) <= 1e-4
    assert nrel(a.x - sc.x, b.x - sc.x) <= 1e-4


@pytest.mark.parametrize("precision", [_lib.SIMP_FP64, _lib.SIMP_MIXED])
def test_rheology_and_kernels_per_instance_in_an_ensemble(precision):
    say(f"ensemble with default, MC, C2, C4, Gaussian, Marquis2 members, precision={precision}")
    scs = S.ensemble(count=6, nmin=1500, nmax=4000, seed=57)
    for sc, sw in zip(scs, ({}, {"rheology": 2}, {"skf": 4}, {"skf": 5}, {"skf": 2},
                            {"skf": 8})):
        sc.time = 1e-6
        for k, v in sw.items():
            setattr(sc, k, v)
    ens = Ensemble(precision)
    ids = ens.add(scs)
    ens.step(6)
    for sc, i in zip(scs, ids):
        got = ens.download(i)
        dev = Simp.for_scenario(sc, precision=precision)
        dev.upload(sc)
        dev.step(6)
        want = dev.download()
        dev.close()
        for k in FIELDS:
            assert np.array_equal(got[k], want[k]), (sc.rheology, sc.skf, k)
        if precision == _lib.SIMP_FP64:
            b = sc.copy()
            capi.Oracle().run_pc(b, 6)
            for k in FIELDS:
                assert nrel(got[k], getattr(b, k)) <= 1e-12, (sc.rheology, sc.skf, k)
    ens.close()
