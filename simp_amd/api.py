"""Host-side mirror of the reference's interface for the SPH-VP hot path.

The reference exposes argument-less Fortran subroutines that work on module
state (``modules/data_variables.f90``).  Here the module state is a
:class:`simp_amd.scenarios.Scenario` and each routine takes it explicitly:

* :meth:`Simp.nnps`             -- ``NNPS`` / ``bucket_search``
  (``neighbor/nnps.f90:21-30,155-517``), pair lists in the reference order.
* :meth:`Simp.single_step`      -- ``single_step``
  (``integration/single_step.f90:5-108``).
* :meth:`Simp.time_integration` -- ``time_integration`` with the
  ``predictor_corrector`` scheme (``integration/time_integration.f90:38-73,
  155-282``), device-resident.

All work runs in ``libsimp_amd.so`` on the GPU; errors raise
:class:`simp_amd._lib.SimpError` where the reference would ``STOP``.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from ._lib import SimpConfig, check, lib, ptr


def make_config(precision: int = _lib.SIMP_FP64, tse: int = 2, wind=(0.0, float(np.float32(-5.3))),
                current=(0.0, 0.0), max_hsml: float = 0.0, wind_grid: bool = False,
                den_alg: int = 3, pa_sph: int = 1, eos_alg: int = 3, rheology: int = 1,
                skf: int = 6, int_algorithm: int = 2, sle: int = 2, xsph: int = 0,
                art_visc: int = 0, art_visc_alg: int = 1, bfa: int = 1,
                boundary_friction: int = 0, boundary_pressure: int = 1,
                pressure_force: int = 1, wind_force: int = 1,
                water_force: int = 1, viscosity: int = 1) -> SimpConfig:
    cfg = SimpConfig()
    lib().simp_config_default(C.byref(cfg))
    cfg.precision = precision
    cfg.boundary_friction = boundary_friction
    cfg.boundary_pressure = boundary_pressure
    cfg.pressure_force = pressure_force
    cfg.wind_force = wind_force
    cfg.water_force = water_force
    cfg.viscosity = viscosity
    cfg.sle = sle
    cfg.xsph_regularization = xsph
    cfg.artificial_visc = art_visc
    cfg.art_visc_alg = art_visc_alg
    cfg.bfa = bfa
    cfg.tse = tse
    cfg.den_alg = den_alg
    cfg.pa_sph = pa_sph
    cfg.eos_alg = eos_alg
    cfg.rheology = rheology
    cfg.skf = skf
    cfg.int_algorithm = int_algorithm
    cfg.wind[0], cfg.wind[1] = wind
    cfg.current[0], cfg.current[1] = current
    cfg.max_hsml = max_hsml
    cfg.wind_mode = _lib.SIMP_WIND_GRID if wind_grid else _lib.SIMP_WIND_UNIFORM
    return cfg


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def _state_arrays(sc):
    """Contiguous FP64 views of a scenario's state, kept alive by the caller
    for the duration of the C call."""
    return [_c(sc.x), _c(sc.vx), _c(sc.mass), _c(sc.rho), _c(sc.hsml), _c(sc.h), _c(sc.A),
            _c(sc.xb), _c(sc.massb)]


def _stats(call) -> dict:
    mx, mn, nz, it = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    avg, nip = C.c_double(), C.c_longlong()
    check(call(C.byref(mx), C.byref(mn), C.byref(avg), C.byref(nz), C.byref(nip), C.byref(it)),
          "simp_interaction_stats")
    return {"max": mx.value, "min": mn.value, "average": avg.value, "no_interaction": nz.value,
            "nip": nip.value, "itimestep": it.value}


class Simp:
    """One simulation context on the current GPU."""

    def __init__(self, precision: int = _lib.SIMP_FP64, tse: int = 2,
                 wind=(0.0, float(np.float32(-5.3))), current=(0.0, 0.0),
                 max_hsml: float = 0.0, wind_grid=None, den_alg: int = 3, pa_sph: int = 1,
                 eos_alg: int = 3, rheology: int = 1, skf: int = 6, int_algorithm: int = 2,
                 sle: int = 2, xsph: int = 0, art_visc: int = 0, art_visc_alg: int = 1,
                 bfa: int = 1, boundary_friction: int = 0, boundary_pressure: int = 1,
                 pressure_force: int = 1, wind_force: int = 1, water_force: int = 1,
                 viscosity: int = 1):
        self.cfg = make_config(precision, tse, wind, current, max_hsml, wind_grid is not None,
                               den_alg, pa_sph, eos_alg, rheology, skf, int_algorithm, sle,
                               xsph, art_visc, art_visc_alg, bfa, boundary_friction,
                               boundary_pressure, pressure_force, wind_force, water_force,
                               viscosity)
        h = C.c_void_p()
        check(lib().simp_create(C.byref(self.cfg), C.byref(h)), "simp_create")
        self._h = h
        self._keep = []
        if wind_grid is not None:
            u, v = _c(wind_grid.u), _c(wind_grid.v)
            self._keep = [u, v]
            check(lib().simp_set_wind_grid(h, u.shape[1], u.shape[0], wind_grid.x0, wind_grid.y0,
                                           wind_grid.dx, wind_grid.dy, ptr(u), ptr(v)),
                  "simp_set_wind_grid")

    @classmethod
    def for_scenario(cls, sc, precision: int = _lib.SIMP_FP64) -> "Simp":
        return cls(precision=precision, tse=sc.tse, wind=sc.wind, current=sc.current,
                   max_hsml=sc.max_hsml, wind_grid=sc.wind_grid,
                   den_alg=sc.den_alg, pa_sph=sc.pa_sph, eos_alg=sc.eos_alg,
                   rheology=sc.rheology, skf=sc.skf, int_algorithm=sc.int_algorithm,
                   sle=sc.sle, xsph=sc.xsph, art_visc=sc.art_visc,
                   art_visc_alg=sc.art_visc_alg, bfa=sc.bfa,
                   boundary_friction=sc.boundary_friction,
                   boundary_pressure=sc.boundary_pressure, pressure_force=sc.pressure_force,
                   wind_force=sc.wind_force, water_force=sc.water_force,
                   viscosity=getattr(sc, "viscosity", 1))

    def set_module_state(self, sc) -> None:
        """dt and the boundary normals of ``sc``, which the Wang boundary
        force reads (simp_set_module_state)."""
        bn = _c(sc.b_normal())
        self._bn = bn
        check(lib().simp_set_module_state(self._h, sc.dt, sc.nb, ptr(bn) if sc.nb else None),
              "simp_set_module_state")

    def close(self):
        if getattr(self, "_h", None):
            lib().simp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        return lib().simp_stream(self._h)

    # ---------------------------------------------------------------- NNPS
    def nnps(self, x, hsml, xb, cap: Optional[int] = None, bcap: Optional[int] = None) -> dict:
        x, hsml, xb = _c(x), _c(hsml), _c(xb)
        n, nb = x.shape[0], xb.shape[0]
        cap = cap if cap is not None else 16 * n + 64
        bcap = bcap if bcap is not None else 64 * max(nb, 1) + 1024
        nip, bnip = C.c_int(), C.c_int()
        pi = np.zeros(cap, np.int32); pj = np.zeros(cap, np.int32)
        r = np.zeros(cap); hij = np.zeros(cap)
        bi = np.zeros(bcap, np.int32); bj = np.zeros(bcap, np.int32); br = np.zeros(bcap)
        st = lib().simp_nnps(self._h, n, ptr(x), ptr(hsml), nb, ptr(xb) if nb else None, cap,
                             C.byref(nip), ptr(pi), ptr(pj), ptr(r), ptr(hij), bcap,
                             C.byref(bnip), ptr(bi), ptr(bj), ptr(br))
        if st == _lib.SIMP_E_CAPACITY and (nip.value > cap or bnip.value > bcap):
            return self.nnps(x, hsml, xb, cap=max(cap, nip.value), bcap=max(bcap, bnip.value))
        check(st, "simp_nnps")
        k, kb = nip.value, bnip.value
        return {"nip": k, "p_i": pi[:k], "p_j": pj[:k], "r_ij": r[:k], "hsml_ij": hij[:k],
                "bnip": kb, "bp_i": bi[:kb], "bp_j": bj[:kb], "br_ij": br[:kb]}

    # --------------------------------------------------------- single_step
    def single_step(self, sc, current_time: Optional[float] = None, itimestep: int = 1,
                    out: Optional[dict] = None, noalloc: bool = False) -> dict:
        """``out``: the output arrays of an earlier call of the same size, to
        be filled again (as a Fortran caller passes its module arrays).
        ``noalloc``: through ``simp_single_step_noalloc``, which allocates
        nothing (its device state lives in the library's per-GPU array)."""
        n = sc.n
        t = sc.time if current_time is None else current_time
        if out is None or out["dhdt"].shape[0] != n:
            out = {k: np.zeros((n, 2)) for k in ("dxdt", "dvxdt")}
            out.update({k: np.zeros(n) for k in ("dhdt", "dAdt", "zeta", "eta", "p")})
            out["eps"] = np.zeros((n, 3))
            out["sigma"] = np.zeros((n, 3))
        nip, bnip = C.c_int(), C.c_int()
        a = _state_arrays(sc)
        self.set_module_state(sc)
        fn = lib().simp_single_step_noalloc if noalloc else lib().simp_single_step
        st = fn(
            self._h, n, ptr(a[0]), ptr(a[1]), ptr(a[2]), ptr(a[3]), ptr(a[4]), ptr(a[5]),
            ptr(a[6]), sc.nb, ptr(a[7]) if sc.nb else None, ptr(a[8]) if sc.nb else None,
            t, itimestep,
            ptr(out["dxdt"]), ptr(out["dvxdt"]), ptr(out["dhdt"]), ptr(out["dAdt"]),
            ptr(out["eps"]), ptr(out["zeta"]), ptr(out["eta"]), ptr(out["p"]),
            ptr(out["sigma"]), C.byref(nip), C.byref(bnip))
        check(st, "simp_single_step_noalloc" if noalloc else "simp_single_step")
        out["nip"], out["bnip"] = nip.value, bnip.value
        return out

    def set_tuning(self, passA_batch: int = 2, passB_batch: int = 0, skin: float = 0.01) -> None:
        """Kernel tuning parameters (no result depends on them)."""
        t = _lib.SimpTuning(passA_batch, passB_batch, 0, skin)
        check(lib().simp_set_tuning(self._h, C.byref(t)), "simp_set_tuning")

    def get_tuning(self) -> dict:
        t = _lib.SimpTuning()
        check(lib().simp_get_tuning(self._h, C.byref(t)), "simp_get_tuning")
        return {"passA_batch": t.passA_batch, "passB_batch": t.passB_batch, "skin": t.skin}

    def tune(self, steps_per_trial: int = 30) -> dict:
        """The autotuner on the uploaded state (its trial steps advance the
        simulation): the setting it keeps and the microseconds per step of
        the setting it started from and of the kept one."""
        t = _lib.SimpTuning()
        a, b = C.c_double(), C.c_double()
        check(lib().simp_tune(self._h, steps_per_trial, C.byref(t), C.byref(a), C.byref(b)),
              "simp_tune")
        return {"passA_batch": t.passA_batch, "passB_batch": t.passB_batch, "skin": t.skin,
                "us_start": a.value, "us_best": b.value}

    def set_search(self, every_evaluation: bool) -> None:
        """Rebuild the cell and candidate lists at every evaluation (as the
        reference's bucket_search does) instead of when needed; the results
        are the same."""
        check(lib().simp_set_search(self._h, 1 if every_evaluation else 0), "simp_set_search")

    def search_stats(self) -> dict:
        """Evaluations since the upload, and of them those that rebuilt the
        cell lists and the candidate lists."""
        e, c, v = C.c_longlong(), C.c_longlong(), C.c_longlong()
        check(lib().simp_search_stats(self._h, C.byref(e), C.byref(c), C.byref(v)),
              "simp_search_stats")
        return {"evaluations": e.value, "cell_rebuilds": c.value,
                "candidate_rebuilds": v.value}

    def interaction_stats(self) -> dict:
        """``interaction_stats`` (``neighbor/nnps.f90:613-658``) of the last
        evaluation that collected it: every ``single_step`` and both
        evaluations of a step that is a multiple of print_step (2000)."""
        return _stats(lambda *a: lib().simp_interaction_stats(self._h, *a))

    def kernel_values(self, n: int, cap: Optional[int] = None) -> dict:
        """The last single_step's pairs with their kernel values w and the
        per-particle kernel_zero (what sum_density reads on the host)."""
        cap = cap if cap is not None else 16 * n + 64
        nip = C.c_int()
        pi = np.zeros(cap, np.int32); pj = np.zeros(cap, np.int32); w = np.zeros(cap)
        kz = np.zeros(n)
        st = lib().simp_kernel_values(self._h, cap, C.byref(nip), ptr(pi), ptr(pj), ptr(w),
                                      ptr(kz) if self.cfg.den_alg == 1 else None)
        if st == _lib.SIMP_E_CAPACITY and nip.value > cap:
            return self.kernel_values(n, cap=nip.value)
        check(st, "simp_kernel_values")
        k = nip.value
        out = {"nip": k, "p_i": pi[:k], "p_j": pj[:k], "w": w[:k]}
        if self.cfg.den_alg == 1:
            out["kernel_zero"] = kz
        return out

    def kernel_gradients(self, cap: Optional[int] = None) -> dict:
        """The last single_step's pairs with their kernel gradients dwdx
        (nip, 2) in the pair's orientation (what con_density reads on the
        host)."""
        cap = cap if cap is not None else 16 * max(getattr(self, "_n", 0), 4096) + 64
        nip = C.c_int()
        pi = np.zeros(cap, np.int32); pj = np.zeros(cap, np.int32); g = np.zeros((cap, 2))
        st = lib().simp_kernel_gradients(self._h, cap, C.byref(nip), ptr(pi), ptr(pj), ptr(g))
        if st == _lib.SIMP_E_CAPACITY and nip.value > cap:
            return self.kernel_gradients(cap=nip.value)
        check(st, "simp_kernel_gradients")
        k = nip.value
        return {"nip": k, "p_i": pi[:k], "p_j": pj[:k], "dwdx": g[:k]}

    # --------------------------------------------- device-resident stepping
    def upload(self, sc, itimestep: int = 0) -> None:
        self._n = sc.n
        a = _state_arrays(sc)
        self.set_module_state(sc)
        check(lib().simp_upload(
            self._h, sc.n, ptr(a[0]), ptr(a[1]), ptr(a[2]), ptr(a[3]), ptr(a[4]), ptr(a[5]),
            ptr(a[6]), sc.nb, ptr(a[7]) if sc.nb else None, ptr(a[8]) if sc.nb else None,
            sc.dt, sc.time, itimestep), "simp_upload")

    def set_diagnostics(self, enable: bool) -> None:
        check(lib().simp_set_diagnostics(self._h, 1 if enable else 0), "simp_set_diagnostics")

    def step(self, nsteps: int, sync: bool = True) -> None:
        check(lib().simp_step_pc(self._h, nsteps, 1 if sync else 0), "simp_step_pc")

    def sync(self) -> None:
        check(lib().simp_sync(self._h), "simp_sync")

    def download(self, sc=None) -> dict:
        n = self._n
        out = {"x": np.zeros((n, 2)), "vx": np.zeros((n, 2)), "rho": np.zeros(n),
               "hsml": np.zeros(n), "h": np.zeros(n), "A": np.zeros(n)}
        dt, t, it = C.c_double(), C.c_double(), C.c_int()
        check(lib().simp_download(self._h, ptr(out["x"]), ptr(out["vx"]), ptr(out["rho"]),
                                  ptr(out["hsml"]), ptr(out["h"]), ptr(out["A"]),
                                  C.byref(dt), C.byref(t), C.byref(it)), "simp_download")
        out.update(dt=dt.value, time=t.value, itimestep=it.value)
        if sc is not None:
            sc.x[:], sc.vx[:], sc.rho[:] = out["x"], out["vx"], out["rho"]
            sc.hsml[:], sc.h[:], sc.A[:] = out["hsml"], out["h"], out["A"]
            sc.dt, sc.time = out["dt"], out["time"]
        return out

    def snapshot(self) -> bytes:
        """Exact snapshot of the device-resident run (``simp_snapshot_save``):
        restored with :meth:`restore` on a handle of the same configuration
        it continues bit for bit."""
        size = lib().simp_snapshot_size(self._h)
        if size < 0:
            check(int(-size), "simp_snapshot_size")
        buf = C.create_string_buffer(size)
        check(lib().simp_snapshot_save(self._h, C.cast(buf, C.c_void_p), size),
              "simp_snapshot_save")
        return buf.raw

    def restore(self, blob: bytes) -> None:
        """Upload a run from :meth:`snapshot` (``simp_snapshot_load``)."""
        check(lib().simp_snapshot_load(self._h, C.c_char_p(blob), len(blob)),
              "simp_snapshot_load")
        self._n = int(np.frombuffer(blob[8:12], dtype="<i4")[0])

    def download_diagnostics(self) -> dict:
        n = self._n
        out = {"eps": np.zeros((n, 3)), "zeta": np.zeros(n), "eta": np.zeros(n),
               "p": np.zeros(n), "sigma": np.zeros((n, 3))}
        check(lib().simp_download_diagnostics(self._h, ptr(out["eps"]), ptr(out["zeta"]),
                                              ptr(out["eta"]), ptr(out["p"]),
                                              ptr(out["sigma"])), "simp_download_diagnostics")
        return out

    def time_integration(self, sc, nsteps: int, itimestep: int = 0) -> dict:
        """Upload ``sc``, advance ``nsteps`` predictor-corrector steps on the
        device and write the result back into ``sc``."""
        self.upload(sc, itimestep)
        self.step(nsteps)
        return self.download(sc)

    def profile_steps(self, nsteps: int) -> dict:
        """Average duration (ms) per launch of each kernel over ``nsteps``
        steps, timed with HIP events on the handle's stream."""
        maxk = 64
        buf = C.create_string_buffer(4096)
        ms = np.zeros(maxk)
        cnt = np.zeros(maxk, np.int32)
        k = lib().simp_profile_steps(self._h, nsteps, maxk, buf, 4096, ptr(ms), ptr(cnt))
        if k < 0:
            check(-k, "simp_profile_steps")
        names = buf.value.decode().split("\n")[:k]
        return {nm: (float(ms[i]), int(cnt[i])) for i, nm in enumerate(names)}
