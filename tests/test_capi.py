"""The C ABI library: it loads, exports every symbol include/simp.h declares,
and validates its inputs on the host (no GPU work here)."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

import pytest

from simp_amd import _lib
from simp_amd.api import make_config

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            text = open(os.path.join(ROOT, "include", f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(simp_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_symbols()
    assert "simp_single_step" in names and "simp_nnps" in names
    for nm in sorted(names):
        assert hasattr(L, nm), nm
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (simp_[a-z0-9_]+)", out))
    assert names <= exported, names - exported


def test_config_defaults_are_the_reference_switches():
    cfg = make_config()
    assert (cfg.int_algorithm, cfg.nnps_algorithm, cfg.sle, cfg.tse) == (2, 2, 2, 2)
    assert (cfg.rheology, cfg.bfa, cfg.eos_alg, cfg.pa_sph) == (1, 1, 3, 1)
    assert (cfg.skf, cfg.den_alg) == (6, 3)
    import numpy as np
    assert cfg.wind[1] == float(np.float32(-5.3))


@pytest.mark.parametrize("field,value", [("skf", 0), ("skf", 9), ("rheology", 3), ("tse", 3),
                                         ("den_alg", 4), ("int_algorithm", 0),
                                         ("int_algorithm", 4),
                                         ("precision", 7), ("pa_sph", 3), ("eos_alg", 0),
                                         ("eos_alg", 4)])
def test_unsupported_switches_are_rejected_before_any_device_work(field, value):
    cfg = make_config()
    setattr(cfg, field, value)
    h = C.c_void_p()
    st = _lib.lib().simp_create(C.byref(cfg), C.byref(h))
    assert st == _lib.SIMP_E_UNSUPPORTED
    assert not h.value


def test_bad_arguments():
    L = _lib.lib()
    assert L.simp_create(None, None) == _lib.SIMP_E_ARG
    cfg = make_config()
    cfg.version = 99
    h = C.c_void_p()
    assert L.simp_create(C.byref(cfg), C.byref(h)) == _lib.SIMP_E_ARG
    assert L.simp_step_pc(None, 1, 1) == _lib.SIMP_E_ARG
    assert L.simp_strerror(_lib.SIMP_E_CAPACITY) == b"capacity exceeded"


def test_tuning_calls_reject_bad_arguments_without_a_device():
    L = _lib.lib()
    t = _lib.SimpTuning(2, 0, 0, 0.01)
    assert L.simp_set_tuning(None, C.byref(t)) == _lib.SIMP_E_ARG
    assert L.simp_get_tuning(None, C.byref(t)) == _lib.SIMP_E_ARG
    assert L.simp_tune(None, 10, None, None, None) == _lib.SIMP_E_ARG
    assert L.simp_batch_tune(None, 10, None, None, None) == _lib.SIMP_E_ARG


def _snapshot_blob(n, nb, nq, itimestep=7, t=2.5, dt=1e-4, pad_doubles=0):
    """A blob in the exact-snapshot layout (magic, int32 n, nb, itimestep,
    pad, int64 nq, dt, time, then 10 n + 3 nb + nq doubles)."""
    import struct
    import numpy as np
    head = b"SIMPSNP1" + struct.pack("<iiiiqdd", n, nb, itimestep, 0, nq, dt, t)
    body = np.arange(10 * max(n, 0) + 3 * max(nb, 0) + max(nq, 0) + pad_doubles,
                     dtype="<f8").tobytes()
    return head + body


def _info(blob, length=None):
    L = _lib.lib()
    n, nb, it, t = C.c_int(-1), C.c_int(-1), C.c_int(-1), C.c_double(-1.0)
    buf = C.create_string_buffer(blob, len(blob))
    st = L.simp_snapshot_info(buf, len(blob) if length is None else length, C.byref(n),
                              C.byref(nb), C.byref(it), C.byref(t))
    return st, n.value, nb.value, it.value, t.value


def test_snapshot_header_is_checked_on_the_host():
    st, n, nb, it, t = _info(_snapshot_blob(5, 3, 11))
    assert (st, n, nb, it, t) == (_lib.SIMP_OK, 5, 3, 7, 2.5)
    good = _snapshot_blob(5, 3, 11)
    # truncated by one byte, or shorter than the header
    assert _info(good[:-1])[0] == _lib.SIMP_E_ARG
    assert _info(good[:40])[0] == _lib.SIMP_E_ARG
    assert _info(good, length=47)[0] == _lib.SIMP_E_ARG
    # not a snapshot
    assert _info(b"SIMPSNP0" + good[8:])[0] == _lib.SIMP_E_ARG
    # forged counts: store sizes whose byte count overflows a signed 64-bit
    # product, negative and zero counts, counts beyond what the blob holds
    for n_, nb_, nq_ in [(5, 3, (1 << 62) + 1), (5, 3, (1 << 61)), (5, 3, -1),
                         (0, 3, 11), (-2, 3, 11), (5, -1, 11), ((1 << 31) - 1, 3, 11),
                         (5, (1 << 31) - 1, 11)]:
        import struct
        forged = good[:8] + struct.pack("<iiiiq", n_, nb_, 7, 0, nq_) + good[32:]
        assert _info(forged)[0] == _lib.SIMP_E_ARG, (n_, nb_, nq_)
    # a longer blob (trailing bytes) is accepted, a null buffer is not
    assert _info(_snapshot_blob(5, 3, 11, pad_doubles=4))[0] == _lib.SIMP_OK
    assert _lib.lib().simp_snapshot_info(None, 100, None, None, None, None) == _lib.SIMP_E_ARG
    # simp_snapshot_load needs a handle: without one it is an argument error
    assert _lib.lib().simp_snapshot_load(None, C.create_string_buffer(good, len(good)),
                                         len(good)) == _lib.SIMP_E_ARG
