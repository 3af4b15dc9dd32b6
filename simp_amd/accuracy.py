"""Per-field error of the mixed-precision path against FP64.

The mixed path runs the kernel-gradient, velocity-gradient, divergence and
stress pair sums in FP32 with compensated (TwoSum) accumulation and keeps
FP64 for positions, the neighbour test, the rheology epilogue (the
Delta -> Delta_min switch), the predictor-corrector update and the
thickness / concentration / density bookkeeping.  This module states what
that costs, field by field, after a fixed number of steps:

* ``norm``  -- normwise relative error ``max|a - b| / max|b|`` of the field;
* ``elem``  -- largest per-element relative error ``|a_i - b_i| /
  (|b_i| + floor)`` with the absolute floor ``1e-3 max|b|``;
* ``delta`` -- for fields that start from a known value (rho, h, A and the
  velocity, which starts at rest): the normwise error of the *change* over
  the run, ``max|(a - a0) - (b - b0)| / max|b - b0|``.  The change is what the
  pair sums drive, so this is the stricter figure when the field itself
  barely moves in the run.

Fields: the density ``rho``, the velocity ``vx`` (both components), the
thickness ``h``, the concentration ``A``, the three stress components of the
last evaluation, and the displacement ``x - x0`` (positions themselves are
dominated by their initial values).
"""
from __future__ import annotations

from typing import Dict, Optional

import numpy as np

from . import _lib
from .api import Simp

FIELDS = ("rho", "vx", "h", "A", "sigma11", "sigma12", "sigma22", "disp")


def run(sc, steps: int, precision: int) -> Dict[str, np.ndarray]:
    """Advance a copy of ``sc`` by ``steps`` device-resident predictor-
    corrector steps; the final state and the stress of the last evaluation."""
    s = sc.copy()
    dev = Simp.for_scenario(s, precision=precision)
    dev.set_diagnostics(True)
    dev.upload(s)
    dev.step(steps)
    out = dev.download()
    out["sigma"] = dev.download_diagnostics()["sigma"]
    dev.close()
    return out


def _fields(st: dict, sc0) -> Dict[str, np.ndarray]:
    return {"rho": st["rho"], "vx": st["vx"], "h": st["h"], "A": st["A"],
            "sigma11": st["sigma"][:, 0], "sigma12": st["sigma"][:, 1],
            "sigma22": st["sigma"][:, 2], "disp": st["x"] - sc0.x}


def _initial(sc0) -> Dict[str, Optional[np.ndarray]]:
    return {"rho": sc0.rho, "vx": sc0.vx, "h": sc0.h, "A": sc0.A}


def field_errors(mixed: dict, fp64: dict, sc0) -> Dict[str, Dict[str, float]]:
    """Error table of ``mixed`` against ``fp64`` (both from :func:`run` or
    the oracle's equivalent dictionaries) for a run started from ``sc0``."""
    a, b = _fields(mixed, sc0), _fields(fp64, sc0)
    init = _initial(sc0)
    out = {}
    for k in FIELDS:
        fa, fb = np.asarray(a[k], float), np.asarray(b[k], float)
        scale = float(np.max(np.abs(fb)))
        d = np.abs(fa - fb)
        row = {"norm": float(d.max() / scale) if scale > 0 else float(d.max()),
               "scale": scale}
        floor = 1e-3 * scale if scale > 0 else 1e-300
        row["elem"] = float(np.max(d / (np.abs(fb) + floor)))
        if k in init:
            ch = np.abs(fb - init[k])
            cs = float(ch.max())
            row["delta"] = float(d.max() / cs) if cs > 0 else float(d.max())
            row["change_scale"] = cs
        out[k] = row
    return out


def report(sc, steps: int) -> Dict[str, Dict[str, float]]:
    """Run ``sc`` for ``steps`` steps in both precisions on the GPU and
    return the per-field error of the mixed run against the FP64 run."""
    m = run(sc, steps, _lib.SIMP_MIXED)
    f = run(sc, steps, _lib.SIMP_FP64)
    return field_errors(m, f, sc)


def worst(table: Dict[str, Dict[str, float]], key: str = "norm") -> float:
    return max(row[key] for row in table.values() if key in row)
