"""Snapshots in the reference's CSV format (``in_out/output.f90``) and the
spin-up reader that reads them back (``in_out/initialize.f90:234-330``),
through the C ABI (``simp_output_*`` of ``include/simp.h``; host code, no
device needed).

A run folder holds ``particle_momentum.csv``, ``particle_state.csv`` and
``particle_other.csv``; every snapshot appends ``n`` records to the first
two and one line to the third, byte for byte as the reference writes them.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import SIMP_E_CAPACITY, check, lib, ptr

_F = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731


def create(folder: str) -> None:
    """Empty (or create) the three files, as the reference's initialize
    does before its first output."""
    os.makedirs(folder, exist_ok=True)
    check(lib().simp_output_create(folder.encode()), "simp_output_create")


def append(folder: str, x, vx, mass, p, hsml, A, h, eps, sigma, nb: int,
           current_time: float) -> None:
    """One snapshot: x, vx (n, 2) in m and m/s; eps, sigma (n, 3) as
    (11, 12, 22); current_time in s (written in hours)."""
    x, vx, eps, sigma = _F(x), _F(vx), _F(eps), _F(sigma)
    n = x.shape[0]
    cols = [_F(a) for a in (mass, p, hsml, A, h)]
    for a in cols:
        assert a.shape == (n,)
    assert vx.shape == (n, 2) and eps.shape == (n, 3) and sigma.shape == (n, 3)
    check(lib().simp_output_append(folder.encode(), n, ptr(x), ptr(vx), *[ptr(a) for a in cols],
                                   ptr(eps), ptr(sigma), nb, current_time),
          "simp_output_append")


def read_last(folder: str, cap: int | None = None) -> dict:
    """The last snapshot of a folder: x, vx back in m and m/s, mass, p,
    hsml, A, h, eps and sigma (n, 3), n, nb, current_time in s."""
    n, nb, t = C.c_int(), C.c_int(), C.c_double()
    L = lib()
    if cap is None:   # size query first
        st = L.simp_output_read_last(folder.encode(), 0, C.byref(n), C.byref(nb), C.byref(t),
                                     *([None] * 9))
        if st not in (0, SIMP_E_CAPACITY):
            check(st, "simp_output_read_last")
        cap = n.value
    out = {"x": np.zeros((cap, 2)), "vx": np.zeros((cap, 2)), "mass": np.zeros(cap),
           "p": np.zeros(cap), "hsml": np.zeros(cap), "A": np.zeros(cap), "h": np.zeros(cap),
           "eps": np.zeros((cap, 3)), "sigma": np.zeros((cap, 3))}
    check(L.simp_output_read_last(folder.encode(), cap, C.byref(n), C.byref(nb), C.byref(t),
                                  *[ptr(out[k]) for k in ("x", "vx", "mass", "p", "hsml", "A",
                                                          "h", "eps", "sigma")]),
          "simp_output_read_last")
    m = n.value
    out = {k: v[:m] for k, v in out.items()}
    out.update(n=m, nb=nb.value, current_time=t.value)
    return out


def snapshot(sim, folder: str, mass, nb: int, current_time: float | None = None) -> dict:
    """Append the state of a device-resident run (``simp_amd.api.Simp``
    stepped with diagnostics enabled) as the reference's output does after a
    step: positions, velocities, hsml, A, h after the step; p, strain and
    stress of its corrector evaluation; the (constant) particle masses come
    from the caller.  ``current_time`` defaults to the run's time after the
    step (the reference writes the time before its increment: pass that
    value to match it)."""
    st = sim.download()
    dg = sim.download_diagnostics()
    t = st["time"] if current_time is None else current_time
    append(folder, st["x"], st["vx"], mass, dg["p"], st["hsml"], st["A"], st["h"],
           dg["eps"], dg["sigma"], nb, t)
    return {**st, **dg}
