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


@pytest.mark.parametrize("field,value", [("skf", 0), ("skf", 9), ("rheology", 3), ("tse", 1),
                                         ("den_alg", 2), ("int_algorithm", 0),
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
