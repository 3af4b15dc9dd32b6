"""Write the golden vectors of the reference SIMP hot path.

Runs the compiled reference (``oracle/_ref/simp_ref_harness``, built from the
reference's own Fortran sources by ``oracle/build_ref.sh``) on small seeded
inputs and stores inputs and outputs as ``.npz`` fixtures next to this
script.  The reference has no tests or fixtures of its own; these vectors are
the pin for the oracle.  Regenerate with::

    bash oracle/build_ref.sh && python tests/golden/make_golden.py [main] [den1] [pa2] [eos] [switches] [regularization] [friction]
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import capi, refrun  # noqa: E402
from simp_amd import scenarios as S  # noqa: E402

STATE_KEYS = ("x", "vx", "mass", "rho", "hsml", "h", "A", "xb", "massb")


def _inputs(sc, prefix="in_"):
    d = {prefix + k: np.asarray(getattr(sc, k)) for k in STATE_KEYS}
    d[prefix + "scalars"] = np.array([sc.dt, sc.time, sc.max_hsml])
    if sc.bnormal is not None:
        d[prefix + "bnormal"] = sc.b_normal()
    return d


def _save(name, **arrays):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **arrays)
    print("wrote", path, os.path.getsize(path), "bytes")


def _outputs(r, keys):
    return {"out_" + k: np.asarray(r[k]) for k in keys}


STEP_KEYS = ("p_i", "p_j", "r_ij", "hsml_ij", "bp_i", "bp_j", "br_ij", "dxdt", "dvxdt",
             "dhdt", "dAdt", "eps", "zeta", "eta", "p", "sigma")
STATE_OUT = ("x", "vx", "rho", "hsml", "A", "h", "eps", "zeta", "eta", "p", "sigma")


def ties():
    """Particles on exact half-bin coordinates (NINT ties) at pair distance
    exactly equal to mhsml, plus boundary particles on ties."""
    n1 = 6
    s = 1000.0
    i, j = np.meshgrid(np.arange(n1), np.arange(n1), indexing="xy")
    x = np.stack([(2 * i.ravel() + 1) * s, (2 * j.ravel() + 1) * s], axis=1).astype(float)
    n = x.shape[0]
    hsml = np.full(n, 2 * s)          # bin = 4000; x / bin = k/2 + 1/4 or ties
    hsml[::5] = 2.5 * s
    h = np.linspace(0.8, 1.2, n)
    A = np.ones(n)
    xb = S.walls(0.0, 0.0, 2 * n1 * s, 2 * n1 * s)
    xb = np.round(xb / 2000.0) * 2000.0     # boundary particles on half-bin ties
    xb = np.unique(xb, axis=0)
    sc = S._finish(x, hsml, h, A, xb, None, tse=2, name="ties")
    return sc


def main():
    if not refrun.available():
        raise SystemExit("build the reference first: bash oracle/build_ref.sh")
    # 1. the 2,500-particle closed box: NNPS and one single_step at t > 0
    sc = S.box(nside=50, spacing=2000.0, seed=1)
    sc.vx = np.random.default_rng(101).normal(0.0, 0.1, size=(sc.n, 2))
    sc.time = 10.0
    r = refrun.run(refrun.MODE_SINGLE_STEP, sc)
    _save("box2500_single_step.npz", **_inputs(sc), **_outputs(r, STEP_KEYS),
          out_nip=r["nip"], out_bnip=r["bnip"])

    # 2. the same box, 20 predictor-corrector steps from rest at t = 0
    #    (no forcing in the first step, VP time step, stale hsml_ij reads)
    sc = S.box(nside=50, spacing=2000.0, seed=1)
    K = 20
    probe = sc.copy()
    o = capi.Oracle()
    times = []
    for k in range(K):
        o.run_pc(probe, 1, itimestep=k)
        times.append(probe.time)
    t_end = 0.5 * (times[-2] + times[-1])
    r = refrun.run(refrun.MODE_TIME_INTEGRATION, sc, integration_time=t_end)
    assert r["itimestep"] == K, r["itimestep"]
    _save("box2500_pc20.npz", **_inputs(sc), **_outputs(r, STATE_OUT),
          out_dt=r["dt"], out_time=r["time"], out_itimestep=r["itimestep"],
          integration_time=t_end)

    # 3. open ice edge with variable hsml and one coast wall (reference wind)
    sc = S.ice_edge(nside=60, spacing=1000.0, seed=3)
    sc.wind_grid = None
    sc.wind = S.REFERENCE_WIND
    K = 10
    probe = sc.copy()
    o = capi.Oracle()
    times = []
    for k in range(K):
        o.run_pc(probe, 1, itimestep=k)
        times.append(probe.time)
    t_end = 0.5 * (times[-2] + times[-1])
    r = refrun.run(refrun.MODE_TIME_INTEGRATION, sc, integration_time=t_end)
    assert r["itimestep"] == K, r["itimestep"]
    _save("edge3600_pc10.npz", **_inputs(sc), **_outputs(r, STATE_OUT),
          out_dt=r["dt"], out_time=r["time"], out_itimestep=r["itimestep"],
          integration_time=t_end)

    # 4. NINT ties and d == mhsml pairs
    sc = ties()
    r = refrun.run(refrun.MODE_NNPS, sc)
    _save("ties_nnps.npz", **_inputs(sc),
          **_outputs(r, ("p_i", "p_j", "r_ij", "hsml_ij", "bp_i", "bp_j", "br_ij")),
          out_nip=r["nip"], out_bnip=r["bnip"])

    # 5. two consecutive single_step calls where the second has fewer
    #    particle pairs than boundary pairs: b_Wendland_C6 reads hsml_ij
    #    entries the first call left behind (boundary_kernel.f90:491)
    a = S.box(nside=12, spacing=1000.0, seed=5, hsml_factor=2.2)
    b = a.copy()
    b.hsml = a.hsml * (1.3 / 2.2)
    b.vx = np.random.default_rng(7).normal(0.0, 0.1, size=(b.n, 2))
    a.time = b.time = 3.0
    r = refrun.run(refrun.MODE_TWO_STATES, a, second=b)
    assert r["bnip"] > r["nip"], (r["nip"], r["bnip"])
    _save("stale_two_calls.npz", **_inputs(a, "a_"), **_inputs(b, "in_"),
          **_outputs(r, STEP_KEYS), out_nip=r["nip"], out_bnip=r["bnip"])


def summation_density():
    """6. the ice edge with summation density (den_alg = 1, density.f90:41-100),
    from the harness built with that PARAMETER: three density calls per step
    on the last single_step's kernel values and the stale kernel_zero."""
    sc = S.ice_edge(nside=60, spacing=1000.0, seed=3)
    sc.wind_grid = None
    sc.wind = S.REFERENCE_WIND
    sc.den_alg = 1
    K = 10
    probe = sc.copy()
    o = capi.Oracle()
    times = []
    for k in range(K):
        o.run_pc(probe, 1, itimestep=k)
        times.append(probe.time)
    t_end = 0.5 * (times[-2] + times[-1])
    r = refrun.run(refrun.MODE_TIME_INTEGRATION, sc, integration_time=t_end,
                   harness=refrun.HARNESS_DEN1)
    assert r["itimestep"] == K, r["itimestep"]
    _save("edge3600_pc10_den1.npz", **_inputs(sc), **_outputs(r, STATE_OUT),
          out_dt=r["dt"], out_time=r["time"], out_itimestep=r["itimestep"],
          integration_time=t_end, den_alg=1)


def stale_kernel_zero():
    """7. two consecutive single_step calls, the second on a permuted state
    with a smaller support: particles that are first in no pair of the
    second call keep the kernel_zero the first call wrote (kernel.f90:503)."""
    a = S.box(nside=12, spacing=1000.0, seed=5, hsml_factor=2.2)
    perm = np.random.default_rng(11).permutation(a.n)
    b = a.copy()
    for k in ("x", "vx", "mass", "rho", "hsml", "h", "A"):
        setattr(b, k, np.ascontiguousarray(getattr(a, k)[perm]))
    b.hsml = b.hsml * (1.3 / 2.2)
    b.vx = np.random.default_rng(7).normal(0.0, 0.1, size=(b.n, 2))
    a.time = b.time = 3.0
    r = refrun.run(refrun.MODE_TWO_STATES, a, second=b, harness=refrun.HARNESS_DEN1)
    owners = set(r["p_i"].tolist())
    stale = [i for i in range(b.n) if i + 1 not in owners]
    assert stale and np.all(r["kernel_zero"][stale] != 0)
    _save("stale_kernel_zero.npz", **_inputs(a, "a_"), **_inputs(b, "in_"),
          **_outputs(r, ("p_i", "p_j", "w", "kernel_zero")), out_nip=r["nip"], den_alg=1)


def stress_2():
    """8. the stress_2 particle approximation (pa_sph = 2,
    internal_stress.f90:189-216), from the harness built with that
    PARAMETER: one single_step of the closed box at t > 0 and ten
    predictor-corrector steps of the ice edge."""
    sc = S.box(nside=50, spacing=2000.0, seed=1)
    sc.vx = np.random.default_rng(101).normal(0.0, 0.1, size=(sc.n, 2))
    sc.time = 10.0
    sc.pa_sph = 2
    r = refrun.run(refrun.MODE_SINGLE_STEP, sc, harness=refrun.HARNESS_PA2)
    _save("box2500_single_step_pa2.npz", **_inputs(sc),
          **_outputs(r, ("dxdt", "dvxdt", "dhdt", "dAdt", "sigma")),
          out_nip=r["nip"], out_bnip=r["bnip"], pa_sph=2)
    sc = S.ice_edge(nside=60, spacing=1000.0, seed=3)
    sc.wind_grid = None
    sc.wind = S.REFERENCE_WIND
    sc.pa_sph = 2
    K = 10
    probe = sc.copy()
    o = capi.Oracle()
    times = []
    for k in range(K):
        o.run_pc(probe, 1, itimestep=k)
        times.append(probe.time)
    t_end = 0.5 * (times[-2] + times[-1])
    r = refrun.run(refrun.MODE_TIME_INTEGRATION, sc, integration_time=t_end,
                   harness=refrun.HARNESS_PA2)
    assert r["itimestep"] == K, r["itimestep"]
    _save("edge3600_pc10_pa2.npz", **_inputs(sc), **_outputs(r, STATE_OUT),
          out_dt=r["dt"], out_time=r["time"], out_itimestep=r["itimestep"],
          integration_time=t_end, pa_sph=2)


def pressures():
    """9. the Hibler and Gutfraind pressures (eos_alg = 1, 2; eos.f90:59,
    94-102), from harnesses built with those PARAMETERs: one single_step of
    the closed box with thickness spread around 1 (rho = h rho_ice, so the
    Gutfraind pressure is not zero) for each, and ten predictor-corrector
    steps of the ice edge with each."""
    for eos, exe in ((1, refrun.HARNESS_EOS1), (2, refrun.HARNESS_EOS2)):
        sc = S.box(nside=50, spacing=2000.0, seed=1)
        rng = np.random.default_rng(103)
        sc.vx = rng.normal(0.0, 0.1, size=(sc.n, 2))
        sc.h = rng.uniform(0.9, 1.1, size=sc.n)
        sc.rho = sc.h * 900.0
        sc.time = 10.0
        sc.eos_alg = eos
        r = refrun.run(refrun.MODE_SINGLE_STEP, sc, harness=exe)
        _save(f"box2500_single_step_eos{eos}.npz", **_inputs(sc),
              **_outputs(r, ("dxdt", "dvxdt", "dhdt", "dAdt", "p", "sigma")),
              out_nip=r["nip"], out_bnip=r["bnip"], eos_alg=eos)
    _pc10_edge("eos1", refrun.HARNESS_EOS1, eos_alg=1)
    _pc10_edge("eos2", refrun.HARNESS_EOS2, eos_alg=2)


def _pc10_edge(tag, harness, K=10, **switches):
    """Ten (or K) steps of the 3,600-particle ice edge (reference wind) with
    the harness built for one switch value."""
    sc = S.ice_edge(nside=60, spacing=1000.0, seed=3)
    sc.wind_grid = None
    sc.wind = S.REFERENCE_WIND
    for k, v in switches.items():
        setattr(sc, k, v)
    probe = sc.copy()
    o = capi.Oracle()
    times = []
    for k in range(K):
        o.run_pc(probe, 1, itimestep=k)
        times.append(probe.time)
    t_end = 0.5 * (times[-2] + times[-1])
    r = refrun.run(refrun.MODE_TIME_INTEGRATION, sc, integration_time=t_end, harness=harness)
    assert r["itimestep"] == K, r["itimestep"]
    assert all(np.all(np.isfinite(r[k])) for k in STATE_OUT), tag
    name = f"edge3600_pc{K}_{tag}.npz"
    _save(name, **_inputs(sc), **_outputs(r, STATE_OUT),
          out_dt=r["dt"], out_time=r["time"], out_itimestep=r["itimestep"],
          integration_time=t_end, **switches)


def switches():
    """10. Mohr-Coulomb rheology (rheology = 2), the Wendland C2 and C4
    kernels (skf = 4, 5), the quadratic, Gaussian, quintic-spline and
    Marquis kernels (skf = 1, 2, 3, 7, 8), the Euler and leap-frog schemes
    (int_algorithm = 1, 3) and the Monaghan time step (tse = 1), each from
    the harness built with that PARAMETER:
    one single_step of the closed box at t > 0 with its pairs' kernel values
    (rheology and kernels), and ten steps of the ice edge (all of them)."""
    kernels = [(f"k{k}", refrun.HARNESS_SKF[k], {"skf": k}) for k in (1, 2, 3, 7, 8)]
    for tag, exe, sw in [("mc", refrun.HARNESS_MC, {"rheology": 2}),
                         ("c2", refrun.HARNESS_C2, {"skf": 4}),
                         ("c4", refrun.HARNESS_C4, {"skf": 5})] + kernels:
        sc = S.box(nside=50, spacing=2000.0, seed=1)
        sc.vx = np.random.default_rng(101).normal(0.0, 0.1, size=(sc.n, 2))
        sc.time = 10.0
        for k, v in sw.items():
            setattr(sc, k, v)
        r = refrun.run(refrun.MODE_SINGLE_STEP, sc, harness=exe)
        _save(f"box2500_single_step_{tag}.npz", **_inputs(sc),
              **_outputs(r, ("dxdt", "dvxdt", "dhdt", "dAdt", "eps", "zeta", "eta", "p",
                             "sigma", "w", "kernel_zero")),
              out_nip=r["nip"], out_bnip=r["bnip"], **sw)
        _pc10_edge(tag, exe, **sw)
    _pc10_edge("euler", refrun.HARNESS_EULER, int_algorithm=1)
    _pc10_edge("leap", refrun.HARNESS_LEAP, int_algorithm=3)
    _pc10_edge("tse1", refrun.HARNESS_TSE1, tse=1)
    _pc10_edge("den2", refrun.HARNESS_DEN2, den_alg=2)


def _single_box(tag, harness, **switches):
    """One single_step of the 2,500-particle closed box at t > 0 with random
    velocities (both signs of v . r in its pairs), kernel values included."""
    sc = S.box(nside=50, spacing=2000.0, seed=1)
    sc.vx = np.random.default_rng(101).normal(0.0, 0.1, size=(sc.n, 2))
    sc.time = 10.0
    for k, v in switches.items():
        setattr(sc, k, v)
    r = refrun.run(refrun.MODE_SINGLE_STEP, sc, harness=harness)
    _save(f"box2500_single_step_{tag}.npz", **_inputs(sc),
          **_outputs(r, ("dxdt", "dvxdt", "dhdt", "dAdt", "eps", "zeta", "eta", "p",
                         "sigma", "w", "kernel_zero")),
          out_nip=r["nip"], out_bnip=r["bnip"], **switches)


def regularization():
    """11. smoothing-length evolution kept (sle = 0) and Benz's rate (sle =
    1, also with the Euler scheme and with continuity density), XSPH, the
    Hernquist-Katz and Monaghan artificial viscosities (the latter also with
    the Monaghan time step and its viscous CFL terms) and the Wang and
    Marquis boundary forces, each from the harness built with those
    PARAMETERs: single_step of the closed box (XSPH, viscosities, Wang) and
    steps of the ice edge (all).  Wang's force is a momentum over dt, which
    the state does not survive for long: two steps (the first, at rest and
    without forcing, leaves the velocities zero; the second moves them to
    about 1 m/s; a third would take them to 1e18)."""
    H = refrun.HARNESS_VARIANT
    _pc10_edge("sle0", H["sle0"], sle=0)
    _pc10_edge("sle1", H["sle1"], sle=1)
    _pc10_edge("sle1eu", H["sle1eu"], sle=1, int_algorithm=1)
    _pc10_edge("sle1d2", H["sle1d2"], sle=1, den_alg=2)
    _single_box("xsph", H["xsph"], xsph=1)
    _pc10_edge("xsph", H["xsph"], xsph=1)
    _single_box("avhk", H["avhk"], art_visc=1)
    _pc10_edge("avhk", H["avhk"], art_visc=1)
    _single_box("avmon", H["avmon"], art_visc=1, art_visc_alg=2)
    _pc10_edge("avmon", H["avmon"], art_visc=1, art_visc_alg=2)
    _pc10_edge("avtse1", H["avtse1"], art_visc=1, tse=1)
    _single_box("bfa2", H["bfa2"], bfa=2)
    _pc10_edge("bfa2", H["bfa2"], K=2, bfa=2)
    _pc10_edge("bfa3", H["bfa3"], bfa=3)


def friction():
    """12. boundary friction: bf_Monaghan's tangential term with the
    gradients of six boundary kernels (C6, quadratic, Gaussian, quintic
    spline, C2, C4), Wang's friction cap, Marquis' braking (with the
    boundary pressure off); the boundary pressure off (no boundary pairs);
    wind, water drag and the replacement pressure off -- single steps of the
    closed box and steps of the ice edge, each from the harness built with
    those PARAMETERs."""
    H = refrun.HARNESS_VARIANT
    F = {"boundary_friction": 1}
    for tag, sw in (("fric", {}), ("fric1", {"skf": 1}), ("fricg", {"skf": 2}),
                    ("fricq", {"skf": 3}), ("fric4", {"skf": 4}), ("fric5", {"skf": 5})):
        _single_box(tag, H[tag], **F, **sw)
    _pc10_edge("fric", H["fric"], **F)
    _single_box("bfa2f", H["bfa2f"], bfa=2, **F)
    _pc10_edge("bfa2f", H["bfa2f"], K=2, bfa=2, **F)
    _single_box("bfa3f", H["bfa3f"], bfa=3, boundary_pressure=0, **F)
    _pc10_edge("bfa3f", H["bfa3f"], bfa=3, boundary_pressure=0, **F)
    _single_box("nobp", H["nobp"], boundary_pressure=0)
    _pc10_edge("nobp", H["nobp"], boundary_pressure=0)
    _single_box("noforce", H["noforce"], wind_force=0, water_force=0, pressure_force=0)
    _pc10_edge("noforce", H["noforce"], wind_force=0, water_force=0, pressure_force=0)


def novisc():
    """13. viscosity off: zeta and eta stay zero, p is what eos sets -- with
    the replacement pressure off too p stays zero and the strain rate is
    never set (single_step.f90:80); with the Hibler eos p is the ice
    strength -- a single step of the closed box and steps of the ice edge
    from the harness built with those PARAMETERs each."""
    H = refrun.HARNESS_VARIANT
    _single_box("novisc", H["novisc"], viscosity=0, pressure_force=0)
    _pc10_edge("novisc", H["novisc"], viscosity=0, pressure_force=0)
    _single_box("novisc1", H["novisc1"], viscosity=0, eos_alg=1)
    _pc10_edge("novisc1", H["novisc1"], viscosity=0, eos_alg=1)


if __name__ == "__main__":
    # no argument: every fixture; otherwise the named groups ("main", "den1",
    # "pa2", "eos", "switches", "regularization", "novisc", "friction")
    todo = sys.argv[1:] or ["main", "den1", "pa2", "eos", "switches", "regularization", "novisc",
                            "friction"]
    if "main" in todo:
        main()
    if "den1" in todo:
        summation_density()
        stale_kernel_zero()
    if "pa2" in todo:
        stress_2()
    if "eos" in todo:
        pressures()
    if "switches" in todo:
        switches()
    if "regularization" in todo:
        regularization()
    if "friction" in todo:
        friction()
    if "novisc" in todo:
        novisc()
