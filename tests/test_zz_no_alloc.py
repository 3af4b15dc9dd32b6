"""simp_single_step_noalloc (the entry point that allocates nothing) against
simp_single_step, through the C ABI.

The two entry points run the same kernels on the same inputs, so their
results must agree bit for bit (mixed precision included): two calls in a
row on one handle each, the second with fewer particles (the persistent
hsml_ij store and kernel_zero carry over from the first call in both).  The
library's per-GPU array keeps the last call's contents, so these tests pass
whatever ran on the GPU before them."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from golden_io import load, scenario
from simp_amd import _lib
from simp_amd import scenarios as S
from simp_amd.api import Simp

pytestmark = pytest.mark.gpu

FIELDS = ("dxdt", "dvxdt", "dhdt", "dAdt", "eps", "zeta", "eta", "p", "sigma")


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _pair(case):
    if case == "ice_edge":
        a = S.ice_edge(nside=90, spacing=1000.0, seed=11)
        b = S.ice_edge(nside=70, spacing=1000.0, seed=12)
    else:
        a = S.box(nside=60, spacing=2000.0, seed=13)
        b = S.box(nside=45, spacing=2000.0, seed=14)
    rng = np.random.default_rng(3)
    for sc, t in ((a, 2.0), (b, 3.0)):
        sc.vx = rng.normal(0.0, 0.2, size=(sc.n, 2))
        sc.time = t
    return a, b


@pytest.mark.parametrize("case,precision,den_alg", [
    ("ice_edge", _lib.SIMP_FP64, 3),
    ("ice_edge", _lib.SIMP_MIXED, 3),
    ("box", _lib.SIMP_FP64, 1),
])
def test_noalloc_equals_single_step_two_calls(case, precision, den_alg):
    a, b = _pair(case)
    assert b.n < a.n and a.nb > 0 and b.nb > 0
    for sc in (a, b):
        sc.den_alg = den_alg
    say(f"noalloc vs single_step: {case} n={a.n} then n={b.n}, precision={precision}, "
        f"den_alg={den_alg}")
    ref = Simp.for_scenario(a, precision=precision)
    dev = Simp.for_scenario(a, precision=precision)
    for i, sc in enumerate((a, b)):
        want = ref.single_step(sc, itimestep=i + 1)
        got = dev.single_step(sc, itimestep=i + 1, noalloc=True)
        assert got["nip"] == want["nip"] and got["bnip"] == want["bnip"], i
        for k in FIELDS:
            assert np.array_equal(got[k], want[k]), (i, k)
        if den_alg == 1:
            kv, kw = dev.kernel_values(sc.n), ref.kernel_values(sc.n)
            assert np.array_equal(kv["kernel_zero"], kw["kernel_zero"]), i
            assert np.array_equal(kv["w"], kw["w"]), i
        assert dev.interaction_stats() == ref.interaction_stats(), i
    ref.close()
    dev.close()


def test_noalloc_stale_hsml_ij_golden():
    """The reference's own output for two calls whose second reads hsml_ij
    entries only the first call wrote (bnip > nip)."""
    say("noalloc: stale hsml_ij golden, two calls")
    g = load("stale_two_calls.npz")
    a, b = scenario(g, "a_"), scenario(g, "in_")
    dev = Simp.for_scenario(a)
    dev.single_step(a, noalloc=True)
    out = dev.single_step(b, noalloc=True)
    assert out["bnip"] > out["nip"]
    for k in ("dvxdt", "dhdt", "sigma"):
        assert np.array_equal(out[k], g["out_" + k]), k
    dev.close()


def test_noalloc_limits_and_sharing():
    say("noalloc: an input beyond the array, then two handles sharing it")
    big = S.box(nside=700, spacing=1000.0, seed=15)       # 490,000 particles
    small = S.box(nside=40, spacing=2000.0, seed=5)
    small.time = 1.0
    dev = Simp.for_scenario(small)
    with pytest.raises(_lib.SimpError) as ei:
        dev.single_step(big, noalloc=True)
    assert ei.value.status == _lib.SIMP_E_CAPACITY
    # the rejected call touched nothing: the next one works
    want = Simp.for_scenario(small).single_step(small)
    got = dev.single_step(small, noalloc=True)
    for k in FIELDS:
        assert np.array_equal(got[k], want[k]), k
    # another handle's call takes over the array: the first handle's state
    # can no longer be read back
    other = Simp.for_scenario(small)
    other.single_step(small, noalloc=True)
    with pytest.raises(_lib.SimpError) as ei:
        dev.interaction_stats()
    assert ei.value.status == _lib.SIMP_E_STATE
    other.interaction_stats()
    # a handle holding a layout it allocated keeps to the allocating calls
    alloc = Simp.for_scenario(small)
    alloc.single_step(small)
    with pytest.raises(_lib.SimpError) as ei:
        alloc.single_step(small, noalloc=True)
    assert ei.value.status == _lib.SIMP_E_STATE
    for h in (dev, other, alloc):
        h.close()
