"""Loading of the golden fixtures written by tests/golden/make_golden.py."""
from __future__ import annotations

import os

import numpy as np

from simp_amd.scenarios import REFERENCE_WIND, Scenario

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name), allow_pickle=False))


def scenario(g, prefix="in_") -> Scenario:
    dt, t, max_hsml = g[prefix + "scalars"]
    sc = Scenario(x=g[prefix + "x"].copy(), vx=g[prefix + "vx"].copy(),
                  mass=g[prefix + "mass"].copy(), rho=g[prefix + "rho"].copy(),
                  hsml=g[prefix + "hsml"].copy(), h=g[prefix + "h"].copy(),
                  A=g[prefix + "A"].copy(), xb=g[prefix + "xb"].reshape(-1, 2).copy(),
                  massb=g[prefix + "massb"].copy(), max_hsml=float(max_hsml),
                  wind=REFERENCE_WIND, tse=int(g.get("tse", 2)), dt=float(dt), time=float(t))
    sc.den_alg = int(g.get("den_alg", 3))
    sc.pa_sph = int(g.get("pa_sph", 1))
    sc.eos_alg = int(g.get("eos_alg", 3))
    sc.rheology = int(g.get("rheology", 1))
    sc.skf = int(g.get("skf", 6))
    sc.int_algorithm = int(g.get("int_algorithm", 2))
    sc.sle = int(g.get("sle", 2))
    sc.xsph = int(g.get("xsph", 0))
    sc.art_visc = int(g.get("art_visc", 0))
    sc.art_visc_alg = int(g.get("art_visc_alg", 1))
    sc.bfa = int(g.get("bfa", 1))
    for k, v in (("boundary_friction", 0), ("boundary_pressure", 1), ("pressure_force", 1),
                 ("wind_force", 1), ("water_force", 1), ("viscosity", 1)):
        setattr(sc, k, int(g.get(k, v)))
    if prefix + "bnormal" in g:
        sc.bnormal = g[prefix + "bnormal"].reshape(-1, 2).copy()
    return sc


def run_pc_like_reference(o, sc, integration_time):
    """Oracle steps until current_time reaches integration_time, the loop
    condition of time_integration.f90:38."""
    it = 0
    while sc.time < integration_time:
        out = o.run_pc(sc, 1, itimestep=it)
        it += 1
    return it, out


def reference_maps(name="reference_default_init.npz"):
    """The generator inputs the reference's initialize used (its two maps
    and their geometry, written by tests/golden/make_init_golden.py), and
    the golden arrays."""
    from simp_amd.scenarios import ReferenceMaps
    g = load(name)
    geo = g["geometry"]
    mp = ReferenceMaps(map=g["map"], bmap=g["bmap"], x_mingeom=float(geo[0]),
                       y_maxgeom=float(geo[1]), map_resolution=float(geo[2]),
                       bx_mingeom=float(geo[3]), by_maxgeom=float(geo[4]),
                       bmap_resolution=float(geo[5]), map_factor=int(geo[6]),
                       resolution=float(geo[7]), b_length=float(geo[8]),
                       hsml_multiplicity=int(geo[9]))
    return mp, g
