"""Exact snapshots (simp_snapshot_*): a run saved after some steps and
restored on a fresh handle continues bit for bit as the uninterrupted run,
including the persistent hsml_ij store the boundary kernel reads and, for
summation density, kernel_zero."""
from __future__ import annotations

import numpy as np
import pytest

from golden_io import load, scenario
from simp_amd import _lib
from simp_amd.api import Simp

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,precision", [("box2500_pc20.npz", _lib.SIMP_FP64),
                                            ("edge3600_pc10_den1.npz", _lib.SIMP_FP64),
                                            ("box2500_pc20.npz", _lib.SIMP_MIXED),
                                            # Benz's persistent drhodt, the evolving masses
                                            ("edge3600_pc10_sle1.npz", _lib.SIMP_FP64),
                                            # the Wang force's boundary normals
                                            ("edge3600_pc2_bfa2.npz", _lib.SIMP_FP64)])
def test_restored_run_continues_bit_for_bit(name, precision):
    print(f"[gpu] snapshot/restore {name} precision {precision}", flush=True)
    sc = scenario(load(name))
    # (the Wang force is a momentum over dt, which the state survives for
    # two steps only: 1 + 1)
    k1, k2 = (1, 1) if sc.bfa == 2 else (4, 6)
    a = Simp.for_scenario(sc, precision=precision)
    a.upload(sc)
    a.step(k1)
    blob = a.snapshot()
    a.step(k2)
    ra = a.download()
    b = Simp.for_scenario(sc, precision=precision)
    b.restore(blob)
    b.step(k2)
    rb = b.download()
    for k in ("x", "vx", "rho", "hsml", "h", "A"):
        assert np.array_equal(ra[k], rb[k]), k
    assert (ra["time"], ra["dt"], ra["itimestep"]) == (rb["time"], rb["dt"], rb["itimestep"])
    assert ra["itimestep"] == k1 + k2


def test_snapshot_holds_the_uploaded_state():
    print("[gpu] snapshot of an uploaded state", flush=True)
    sc = scenario(load("box2500_pc20.npz"))
    a = Simp.for_scenario(sc)
    a.upload(sc, itimestep=7)
    blob = a.snapshot()
    n, nb = sc.n, sc.nb
    assert len(blob) == 48 + 8 * (10 * n + 3 * nb + int(np.frombuffer(blob[24:32], "<i8")[0]))
    p = np.frombuffer(blob[48:], dtype="<f8")
    assert np.array_equal(p[:2 * n].reshape(n, 2), sc.x)
    assert np.array_equal(p[4 * n:5 * n], sc.mass)
    assert np.array_equal(p[9 * n:9 * n + 2 * nb].reshape(nb, 2), sc.xb)
    assert int(np.frombuffer(blob[16:20], "<i4")[0]) == 7
    b = Simp.for_scenario(sc)
    with pytest.raises(_lib.SimpError):
        b.restore(b"not a snapshot" + blob[14:])
