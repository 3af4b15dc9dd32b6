"""Many independent simulations stepped together in one packed launch.

The reference runs one simulation per process (``modules/globals.f90:21-25``);
an :class:`Ensemble` packs instances of uneven size, domain and forcing into
one set of device buffers with per-instance offsets, so every kernel of a
step covers all of them.  Instances can be added and removed between steps
(the layout is rebuilt on the device); each instance carries its own step
budget and freezes when it is used up or when it fails, without affecting
the others.  Results are identical to running each instance alone (the
kernels never mix instances; see ``tests/test_gpu_ensemble.py``).
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import SimpInstanceDesc, check, lib


def _desc(sc, nsteps: int, keep: list, itimestep: int = 0) -> SimpInstanceDesc:
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in
            (sc.x, sc.vx, sc.mass, sc.rho, sc.hsml, sc.h, sc.A, sc.xb, sc.massb)]
    keep.extend(arrs)
    d = SimpInstanceDesc()
    d.n, d.nb = sc.n, sc.nb
    (d.x, d.vx, d.mass, d.rho, d.hsml, d.hth, d.A, d.xb, d.massb) = [
        a.ctypes.data if a.size else None for a in arrs]
    d.dt, d.current_time = sc.dt, sc.time
    d.itimestep = int(itimestep)
    d.nsteps = int(nsteps)
    d.tse = sc.tse
    d.den_alg = getattr(sc, "den_alg", 3)
    d.pa_sph = getattr(sc, "pa_sph", 1)
    d.eos_alg = getattr(sc, "eos_alg", 3)
    d.rheology = getattr(sc, "rheology", 1)
    d.skf = getattr(sc, "skf", 6)
    sle = getattr(sc, "sle", 2)
    d.sle = _lib.SIMP_SLE_KEEP if sle == 0 else sle
    d.xsph = getattr(sc, "xsph", 0)
    d.art_visc = getattr(sc, "art_visc", 0)
    d.art_visc_alg = getattr(sc, "art_visc_alg", 1)
    d.bfa = getattr(sc, "bfa", 1)
    d.boundary_friction = getattr(sc, "boundary_friction", 0)
    off = 0
    for flag, name in ((_lib.SIMP_OFF_BOUNDARY_PRESSURE, "boundary_pressure"),
                       (_lib.SIMP_OFF_PRESSURE_FORCE, "pressure_force"),
                       (_lib.SIMP_OFF_WIND, "wind_force"), (_lib.SIMP_OFF_WATER, "water_force"),
                       (_lib.SIMP_OFF_VISCOSITY, "viscosity")):
        if not getattr(sc, name, 1):
            off |= flag
    d.forces_off = off
    if sc.bnormal is not None and sc.nb:
        bn = np.ascontiguousarray(sc.b_normal(), dtype=np.float64)
        keep.append(bn)
        d.b_normal = bn.ctypes.data
    d.wind[0], d.wind[1] = sc.wind
    d.current[0], d.current[1] = sc.current
    d.max_hsml = sc.max_hsml
    if sc.wind_grid is not None:
        g = sc.wind_grid
        u = np.ascontiguousarray(g.u, dtype=np.float64)
        v = np.ascontiguousarray(g.v, dtype=np.float64)
        keep.extend([u, v])
        d.wind_mode = _lib.SIMP_WIND_GRID
        d.grid_ny, d.grid_nx = u.shape
        d.grid_x0, d.grid_y0, d.grid_dx, d.grid_dy = g.x0, g.y0, g.dx, g.dy
        d.grid_u, d.grid_v = u.ctypes.data, v.ctypes.data
    else:
        d.wind_mode = _lib.SIMP_WIND_UNIFORM
    return d


def footprint(sc) -> int:
    """Device bytes the scenario would occupy inside an ensemble."""
    keep: list = []
    d = _desc(sc, 0, keep)
    return int(lib().simp_instance_footprint(C.byref(d)))


class Ensemble:
    def __init__(self, precision: int = _lib.SIMP_MIXED):
        h = C.c_void_p()
        check(lib().simp_batch_create(precision, C.byref(h)), "simp_batch_create")
        self._h = h
        self._n = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().simp_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, scenarios: Sequence, nsteps: Optional[Iterable[int]] = None,
            itimestep: int = 0) -> List[int]:
        """Add instances (each may carry a step budget); their step counters
        start at ``itimestep``."""
        scs = list(scenarios)
        ns = list(nsteps) if nsteps is not None else [0] * len(scs)
        keep: list = []
        arr = (SimpInstanceDesc * len(scs))(*[_desc(sc, k, keep, itimestep)
                                              for sc, k in zip(scs, ns)])
        ids = np.zeros(len(scs), np.int32)
        check(lib().simp_batch_add(self._h, len(scs), arr, ids.ctypes.data), "simp_batch_add")
        for sc, i in zip(scs, ids):
            self._n[int(i)] = sc.n
        return [int(i) for i in ids]

    def remove(self, ids: Iterable[int]) -> None:
        v = np.asarray(list(ids), np.int32)
        check(lib().simp_batch_remove(self._h, len(v), v.ctypes.data), "simp_batch_remove")
        for i in v:
            self._n.pop(int(i), None)

    def update(self, remove: Iterable[int], scenarios: Sequence,
               nsteps: Optional[Iterable[int]] = None) -> List[int]:
        """Remove instances and add new ones with one layout rebuild; the
        ids of the added instances."""
        rm = np.asarray(list(remove), np.int32)
        scs = list(scenarios)
        ns = list(nsteps) if nsteps is not None else [0] * len(scs)
        keep: list = []
        arr = (SimpInstanceDesc * max(len(scs), 1))(*[_desc(sc, k, keep) for sc, k in zip(scs, ns)])
        ids = np.zeros(max(len(scs), 1), np.int32)
        check(lib().simp_batch_update(self._h, len(rm), rm.ctypes.data if len(rm) else None,
                                      len(scs), arr, ids.ctypes.data), "simp_batch_update")
        for i in rm:
            self._n.pop(int(i), None)
        for sc, i in zip(scs, ids):
            self._n[int(i)] = sc.n
        return [int(i) for i in ids[:len(scs)]]

    def step(self, nsteps: int, sync: bool = True) -> None:
        check(lib().simp_batch_step(self._h, nsteps, 1 if sync else 0), "simp_batch_step")

    def sync(self) -> None:
        check(lib().simp_batch_sync(self._h), "simp_batch_sync")

    def tune(self, steps_per_trial: int = 20) -> dict:
        """The autotuner on the packed content (its trial steps are steps of
        every active instance)."""
        from ._lib import SimpTuning
        t = SimpTuning()
        a, b = C.c_double(), C.c_double()
        check(lib().simp_batch_tune(self._h, steps_per_trial, C.byref(t), C.byref(a),
                                    C.byref(b)), "simp_batch_tune")
        return {"passA_batch": t.passA_batch, "passB_batch": t.passB_batch, "skin": t.skin,
                "us_start": a.value, "us_best": b.value}

    def status(self, id_: int) -> dict:
        it, left, act, st = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(lib().simp_batch_status(self._h, id_, C.byref(it), C.byref(left), C.byref(act),
                                      C.byref(st)), "simp_batch_status")
        return {"itimestep": it.value, "steps_left": left.value, "active": bool(act.value),
                "status": st.value}

    def download(self, id_: int, sc=None) -> dict:
        n = self._n[id_]
        out = {"x": np.zeros((n, 2)), "vx": np.zeros((n, 2)), "rho": np.zeros(n),
               "hsml": np.zeros(n), "h": np.zeros(n), "A": np.zeros(n)}
        dt, t, it = C.c_double(), C.c_double(), C.c_int()
        check(lib().simp_batch_download(
            self._h, id_, out["x"].ctypes.data, out["vx"].ctypes.data, out["rho"].ctypes.data,
            out["hsml"].ctypes.data, out["h"].ctypes.data, out["A"].ctypes.data,
            C.byref(dt), C.byref(t), C.byref(it)), "simp_batch_download")
        out.update(dt=dt.value, time=t.value, itimestep=it.value)
        if sc is not None:
            sc.x[:], sc.vx[:], sc.rho[:] = out["x"], out["vx"], out["rho"]
            sc.hsml[:], sc.h[:], sc.A[:] = out["hsml"], out["h"], out["A"]
            sc.dt, sc.time = out["dt"], out["time"]
        return out

    def list(self) -> dict:
        """Status of every instance (one device-to-host copy):
        {id: {itimestep, steps_left, active, status}}."""
        cap = max(len(self._n), 1)
        a = [np.zeros(cap, np.int32) for _ in range(5)]
        k = lib().simp_batch_list(self._h, cap, *[x.ctypes.data for x in a])
        if k < 0:
            check(-k, "simp_batch_list")
        return {int(a[0][j]): {"itimestep": int(a[1][j]), "steps_left": int(a[2][j]),
                               "active": bool(a[3][j]), "status": int(a[4][j])}
                for j in range(min(k, cap))}

    def interaction_stats(self, id_: int) -> dict:
        """Neighbour statistics of one instance (see ``Simp.interaction_stats``)."""
        from .api import _stats
        return _stats(lambda *a: lib().simp_batch_interaction_stats(self._h, id_, *a))

    def count(self) -> tuple:
        ni, npart = C.c_int(), C.c_longlong()
        check(lib().simp_batch_count(self._h, C.byref(ni), C.byref(npart)), "simp_batch_count")
        return ni.value, npart.value
