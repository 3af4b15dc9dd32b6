"""Synthetic initial states for the SPH viscous-plastic step.

Each scenario follows the reference's initialisation conventions
(``in_out/initialize.f90:190-197``): ``rho = rho_ice * h``,
``mass = rho * (hsml / 3)**2``, ``max_hsml = 10 * max(hsml)``, velocities at
rest, and boundary particles with ``massb = mean(mass)``
(``in_out/boundaries.f90:52``) placed on the walls at spacing
``b_length / 3`` (``in_out/boundaries.f90:204``).  Lattice particles sit at
``((i + 1/2) s, (j + 1/2) s)`` with i.i.d. uniform jitter of +-5 % of ``s``.

Named workloads:

* ``box``         closed square box (4 walls), uniform thickness spread.
* ``vortex``      closed box, lognormal thickness, Rankine-vortex wind.
* ``ice_edge``    open ice edge: concentration ramp, variable ``hsml``,
                  one coast wall, Gaussian-random-field wind.
* ``ensemble``    many independent boxes of uneven size.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

RHO_ICE = 900.0
B_LENGTH = 8000.0
HSML_MULT = 3.0
#: the reference's hard-coded wind, a default-REAL literal widened to REAL(8)
#: (``ext_force/wind_forcing.f90:36-37``)
REFERENCE_WIND = (0.0, float(np.float32(-5.3)))


@dataclasses.dataclass
class WindGrid:
    """Static wind field on a regular grid, sampled bilinearly."""
    x0: float
    y0: float
    dx: float
    dy: float
    u: np.ndarray  # (ny, nx)
    v: np.ndarray  # (ny, nx)


@dataclasses.dataclass
class Scenario:
    """Initial state of one simulation (host arrays, FP64)."""
    x: np.ndarray        # (n, 2)
    vx: np.ndarray       # (n, 2)
    mass: np.ndarray     # (n,)
    rho: np.ndarray
    hsml: np.ndarray
    h: np.ndarray
    A: np.ndarray
    xb: np.ndarray       # (nb, 2)
    massb: np.ndarray    # (nb,)
    max_hsml: float
    wind: tuple = REFERENCE_WIND
    wind_grid: Optional[WindGrid] = None
    current: tuple = (0.0, 0.0)
    tse: int = 2
    den_alg: int = 3         # 3 diagnostic density, 1 summation density
    pa_sph: int = 1          # 1 stress_1: sigma/rho^2 sums; 2 stress_2: over rho_i rho_j
    eos_alg: int = 3         # 3 Kreyscher, 1 Hibler P, 2 Gutfraind pressure
    rheology: int = 1        # 1 elliptical yield curve, 2 Mohr-Coulomb
    skf: int = 6             # 6 Wendland C6, 4 Wendland C2, 5 Wendland C4
    int_algorithm: int = 2   # 2 predictor-corrector, 1 Euler, 3 leap-frog
    sle: int = 2             # 2 Marquis hsml, 1 Benz hsml, 0 hsml unchanged
    xsph: int = 0            # 1: XSPH position regularization
    art_visc: int = 0        # 1: artificial viscosity
    art_visc_alg: int = 1    # 1 Hernquist and Katz, 2 Monaghan
    bfa: int = 1             # boundary force: 1 Monaghan, 2 Wang, 3 Marquis
    # the logical scenario switches of globals.f90:178-187 (1 = .TRUE.)
    boundary_friction: int = 0
    boundary_pressure: int = 1
    pressure_force: int = 1
    viscosity: int = 1       # 0 (with pressure_force 0): no strain rate
    wind_force: int = 1
    water_force: int = 1
    #: inward unit normal of each boundary particle (nb, 2), as
    #: boundary_placing_1 sets b_normal (in_out/boundaries.f90:232-287);
    #: read by the Wang boundary force only.  None: zeros.
    bnormal: Optional[np.ndarray] = None
    dt: float = 0.0
    time: float = 0.0
    name: str = ""

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    @property
    def nb(self) -> int:
        return int(self.xb.shape[0])

    def copy(self) -> "Scenario":
        return dataclasses.replace(
            self, x=self.x.copy(), vx=self.vx.copy(), mass=self.mass.copy(),
            rho=self.rho.copy(), hsml=self.hsml.copy(), h=self.h.copy(),
            A=self.A.copy(), xb=self.xb.copy(), massb=self.massb.copy(),
            bnormal=None if self.bnormal is None else self.bnormal.copy())

    def b_normal(self) -> np.ndarray:
        """The boundary normals as (nb, 2) FP64 (zeros when unset)."""
        if self.bnormal is None or self.nb == 0:
            return np.zeros((self.nb, 2))
        return np.ascontiguousarray(self.bnormal, dtype=np.float64).reshape(self.nb, 2)


def vp_dt(hsml: np.ndarray) -> float:
    """VP time step, ``integration/time_step.f90:155``."""
    return float(np.min(RHO_ICE * (hsml * hsml) * 2.0e-9 / (2.75e4 * 1.0)))


def lattice(nx: int, ny: int, s: float, rng: np.random.Generator,
            jitter: float = 0.05, x0: float = 0.0, y0: float = 0.0) -> np.ndarray:
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    pts = np.stack([(i.ravel() + 0.5) * s + x0, (j.ravel() + 0.5) * s + y0], axis=1)
    pts += rng.uniform(-jitter * s, jitter * s, size=pts.shape)
    return pts


def walls(x0: float, y0: float, x1: float, y1: float,
          sides=("left", "right", "bottom", "top")) -> np.ndarray:
    """Boundary particles on the named sides of a rectangle, spacing at most
    ``b_length / 3``; corners are shared, not duplicated."""
    sp = B_LENGTH / 3.0
    pts = []

    def seg(a, b):
        L = math.hypot(b[0] - a[0], b[1] - a[1])
        m = max(1, int(math.ceil(L / sp)))
        t = np.arange(m + 1) / m
        return np.stack([a[0] + t * (b[0] - a[0]), a[1] + t * (b[1] - a[1])], axis=1)

    corners = {"bottom": ((x0, y0), (x1, y0)), "top": ((x0, y1), (x1, y1)),
               "left": ((x0, y0), (x0, y1)), "right": ((x1, y0), (x1, y1))}
    for s_ in sides:
        pts.append(seg(*corners[s_]))
    if not pts:
        return np.zeros((0, 2))
    P = np.concatenate(pts)
    # drop duplicated corners (exact coordinates)
    _, idx = np.unique(np.round(P, 6), axis=0, return_index=True)
    return P[np.sort(idx)]


def wall_normals(xb: np.ndarray, x0: float, y0: float, x1: float, y1: float) -> np.ndarray:
    """Inward unit normals of boundary particles on the walls of the
    rectangle [x0, x1] x [y0, y1] (the orientation boundary_placing_1 gives
    its rows, in_out/boundaries.f90:232-287: pointing into the ice); a
    corner takes the normal of its vertical wall."""
    nrm = np.zeros_like(xb)
    nrm[xb[:, 1] == y0] = (0.0, 1.0)
    nrm[xb[:, 1] == y1] = (0.0, -1.0)
    nrm[xb[:, 0] == x0] = (1.0, 0.0)
    nrm[xb[:, 0] == x1] = (-1.0, 0.0)
    return nrm


def lognormal_thickness(rng, n, mean=1.5, sd=0.5):
    s2 = math.log(1.0 + sd * sd / (mean * mean))
    mu = math.log(mean * mean / math.sqrt(mean * mean + sd * sd))
    return rng.lognormal(mu, math.sqrt(s2), size=n)


def _finish(x, hsml, h, A, xb, rng, **kw) -> Scenario:
    n = x.shape[0]
    rho = RHO_ICE * h
    mass = rho * (hsml / HSML_MULT) ** 2
    massb = np.full(xb.shape[0], mass.sum() / n)
    sc = Scenario(x=np.ascontiguousarray(x, dtype=np.float64),
                  vx=np.zeros((n, 2)), mass=mass, rho=rho,
                  hsml=np.ascontiguousarray(hsml, dtype=np.float64),
                  h=np.ascontiguousarray(h, dtype=np.float64),
                  A=np.ascontiguousarray(A, dtype=np.float64),
                  xb=np.ascontiguousarray(xb, dtype=np.float64), massb=massb,
                  max_hsml=float(hsml.max() * 10.0), **kw)
    if sc.tse == 2 and sc.dt == 0.0:
        sc.dt = vp_dt(sc.hsml)
    return sc


def box(nside: int = 50, spacing: float = 2000.0, seed: int = 1,
        hsml_factor: float = 1.3, wind=REFERENCE_WIND, h_range=(0.8, 1.2),
        tse: int = 2, dt: float = 0.0) -> Scenario:
    """Closed square box with 4 walls (the 2,500-particle case at defaults)."""
    rng = np.random.default_rng(seed)
    L = nside * spacing
    x = lattice(nside, nside, spacing, rng)
    n = x.shape[0]
    hsml = np.full(n, hsml_factor * spacing)
    h = rng.uniform(h_range[0], h_range[1], size=n)
    A = np.ones(n)
    xb = walls(0.0, 0.0, L, L)
    return _finish(x, hsml, h, A, xb, rng, wind=tuple(wind), tse=tse, dt=dt,
                   name=f"box{n}", bnormal=wall_normals(xb, 0.0, 0.0, L, L))


def rankine_grid(L: float, vmax: float, radius: float, gdx: float) -> WindGrid:
    m = int(math.ceil(L / gdx)) + 1
    g = np.arange(m) * gdx
    X, Y = np.meshgrid(g, g, indexing="xy")
    cx = cy = L / 2.0
    rx, ry = X - cx, Y - cy
    r = np.hypot(rx, ry)
    speed = np.where(r <= radius, vmax * r / radius,
                     vmax * radius / np.maximum(r, 1e-9))
    with np.errstate(invalid="ignore", divide="ignore"):
        u = np.where(r > 0, -speed * ry / np.maximum(r, 1e-9), 0.0)
        v = np.where(r > 0, speed * rx / np.maximum(r, 1e-9), 0.0)
    return WindGrid(0.0, 0.0, gdx, gdx, np.ascontiguousarray(u), np.ascontiguousarray(v))


def vortex(nside: int = 200, spacing: float = 1000.0, seed: int = 2,
           tse: int = 2, dt: float = 0.0) -> Scenario:
    """Closed box, lognormal thickness, counter-clockwise Rankine vortex wind
    (15 m/s at 60 km) -- the 40,000-particle case at defaults."""
    rng = np.random.default_rng(seed)
    L = nside * spacing
    x = lattice(nside, nside, spacing, rng)
    n = x.shape[0]
    hsml = np.full(n, 1.3 * spacing)
    h = lognormal_thickness(rng, n)
    A = np.ones(n)
    xb = walls(0.0, 0.0, L, L)
    grid = rankine_grid(L, 15.0, 60000.0 * (L / 200000.0), 2000.0 * (L / 200000.0))
    return _finish(x, hsml, h, A, xb, rng, wind_grid=grid, tse=tse, dt=dt,
                   name=f"vortex{n}", bnormal=wall_normals(xb, 0.0, 0.0, L, L))


def grf_wind(L: float, std: float, corr: float, gdx: float,
             rng: np.random.Generator) -> WindGrid:
    """Gaussian random field per wind component with Gaussian covariance of
    length ``corr``, synthesised by FFT on a ``gdx`` grid."""
    m = int(math.ceil(L / gdx)) + 1
    M = 1
    while M < 2 * m:
        M *= 2
    k = np.fft.fftfreq(M, d=gdx) * 2 * math.pi
    KX, KY = np.meshgrid(k, k, indexing="xy")
    # spectrum of exp(-r^2 / (2 l^2)) in 2-D
    spec = np.exp(-0.5 * (KX ** 2 + KY ** 2) * corr ** 2)
    amp = np.sqrt(spec)
    out = []
    for _ in range(2):
        noise = rng.standard_normal((M, M))
        f = np.real(np.fft.ifft2(np.fft.fft2(noise) * amp))[:m, :m]
        f = (f - f.mean()) / max(f.std(), 1e-30) * std
        out.append(np.ascontiguousarray(f))
    return WindGrid(0.0, 0.0, gdx, gdx, out[0], out[1])


def ice_edge(nside: int = 500, spacing: float = 1000.0, seed: int = 3,
             tse: int = 2, dt: float = 0.0) -> Scenario:
    """Open ice edge: concentration 1 up to 80 % of the domain then a linear
    ramp to 0.2, ``hsml = 1.3 s / sqrt(A)``, one coast wall at x = 0,
    Gaussian-random-field wind (std 8 m/s, 50 km correlation) -- the
    250,000-particle case at defaults."""
    rng = np.random.default_rng(seed)
    L = nside * spacing
    x = lattice(nside, nside, spacing, rng)
    n = x.shape[0]
    xr = x[:, 0] / L
    A = np.where(xr <= 0.8, 1.0, 1.0 - 0.8 * (xr - 0.8) / 0.2)
    A = np.clip(A, 0.2, 1.0)
    hsml = 1.3 * spacing / np.sqrt(A)
    h = lognormal_thickness(rng, n)
    xb = walls(0.0, 0.0, 0.0, L, sides=("left",))
    grid = grf_wind(L, 8.0, 50000.0 * (L / 500000.0), 5000.0 * (L / 500000.0), rng)
    return _finish(x, hsml, h, A, xb, rng, wind_grid=grid, tse=tse, dt=dt,
                   name=f"ice_edge{n}", bnormal=np.tile([1.0, 0.0], (xb.shape[0], 1)))


def box_member(N: float, L: float, wind, seed: int, tse: int = 2) -> Scenario:
    """One ensemble member: a closed box of side L holding about N particles
    (lattice spacing L / sqrt(N), hsml = 1.3 s), lognormal thickness,
    uniform wind."""
    nside = max(2, int(round(math.sqrt(N))))
    s = L / nside
    sub = np.random.default_rng(seed)
    x = lattice(nside, nside, s, sub)
    n = x.shape[0]
    hsml = np.full(n, 1.3 * s)
    h = lognormal_thickness(sub, n)
    A = np.ones(n)
    xb = walls(0.0, 0.0, L, L)
    return _finish(x, hsml, h, A, xb, sub, wind=tuple(wind), tse=tse, name=f"box{n}_L{L:.0f}",
                   bnormal=wall_normals(xb, 0.0, 0.0, L, L))


def ensemble(count: int = 64, nmin: int = 2000, nmax: int = 60000,
             seed: int = 4, tse: int = 2) -> List[Scenario]:
    """Independent closed boxes of lognormal particle count in [nmin, nmax],
    box size U[50, 300] km, uniform wind of random direction, U[5, 20] m/s."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        N = float(np.clip(rng.lognormal(math.log(12000.0), 0.8), nmin, nmax))
        L = rng.uniform(50e3, 300e3)
        ang = rng.uniform(0, 2 * math.pi)
        spd = rng.uniform(5.0, 20.0)
        out.append(box_member(N, L, (spd * math.cos(ang), spd * math.sin(ang)),
                              seed * 100003 + k, tse=tse))
    return out


def subdomain(sc: Scenario, x0: float, x1: float, y0: float, y1: float) -> Scenario:
    """The particles and boundary particles of ``sc`` inside a rectangle (a
    bounded sample of a larger workload)."""
    m = (sc.x[:, 0] >= x0) & (sc.x[:, 0] < x1) & (sc.x[:, 1] >= y0) & (sc.x[:, 1] < y1)
    mb = (sc.xb[:, 0] >= x0 - B_LENGTH) & (sc.xb[:, 0] < x1 + B_LENGTH) & \
         (sc.xb[:, 1] >= y0 - B_LENGTH) & (sc.xb[:, 1] < y1 + B_LENGTH)
    out = dataclasses.replace(
        sc, x=np.ascontiguousarray(sc.x[m]), vx=np.ascontiguousarray(sc.vx[m]),
        mass=sc.mass[m].copy(), rho=sc.rho[m].copy(), hsml=sc.hsml[m].copy(),
        h=sc.h[m].copy(), A=sc.A[m].copy(), xb=np.ascontiguousarray(sc.xb[mb]),
        massb=sc.massb[mb].copy(), name=sc.name + "_sub",
        bnormal=None if sc.bnormal is None else np.ascontiguousarray(sc.bnormal[mb]))
    if out.tse == 2:
        out.dt = vp_dt(out.hsml)
    return out


@dataclasses.dataclass
class ReferenceMaps:
    """The inputs of the reference's scenario setup: the domain map
    (``modules/domain.f90``, 1 = ice cell) and the boundary map
    (``modules/boundary_domain.f90``, 1 = ocean/ice side), both indexed
    (x cell, y cell) with the first row at the top, their geometry, and the
    particle resolution and boundary length of ``modules/globals.f90`` /
    ``data_variables.f90``."""
    map: np.ndarray            # (x_numcell, y_numcell) int
    bmap: np.ndarray           # (bx_numcell, by_numcell) int
    x_mingeom: float
    y_maxgeom: float
    map_resolution: float
    bx_mingeom: float
    by_maxgeom: float
    bmap_resolution: float
    map_factor: int
    resolution: float
    b_length: float = B_LENGTH
    hsml_multiplicity: int = 3


def _f32(v):
    return np.float32(v)


def from_maps(mp: ReferenceMaps, wind=REFERENCE_WIND, tse: int = 2) -> Scenario:
    """The reference's default initial state, built as its initialize does
    (``in_out/initialize.f90:102-210`` particle_placing, then
    ``in_out/boundaries.f90:164-312`` boundary_placing_1 and the boundary
    masses of ``boundaries.f90:46``), with the reference's default-REAL
    (single precision) intermediate arithmetic kept where its literals make
    it so: ``-1./2. + m + (k + 1./2.) / res_ratio`` is evaluated in single
    precision and only then scaled by the (double) map resolution."""
    res_ratio = int(np.rint(mp.map_resolution / mp.resolution))   # NINT
    if res_ratio < 1:
        raise ValueError("particle resolution too coarse for the domain map")
    nx, ny = mp.map.shape
    xs = []
    half = _f32(0.5)
    rr = _f32(res_ratio)
    for n in range(ny):                  # initialize.f90:166-185, y cells outer
        for m in range(nx):
            if mp.map[m, n] != 1:
                continue
            for k in range(res_ratio):
                for l in range(res_ratio):
                    ax = (-half + _f32(m)) + (_f32(k) + half) / rr
                    ay = (half - _f32(n)) - (_f32(l) + half) / rr
                    xs.append((mp.x_mingeom + float(ax) * mp.map_resolution,
                               mp.y_maxgeom + float(ay) * mp.map_resolution))
    x = np.array(xs, dtype=np.float64).reshape(-1, 2)
    n_p = x.shape[0]
    hsml = np.full(n_p, mp.hsml_multiplicity * mp.map_resolution / res_ratio)
    max_hsml = float(np.max(hsml)) * 10.0
    A = np.ones(n_p)
    h = np.ones(n_p)
    rho = RHO_ICE * h
    mass = rho * (hsml / mp.hsml_multiplicity) ** 2

    # boundary_placing_1: one boundary particle row along every edge between
    # an outside cell (0) and an inside cell (1) of the boundary map, bppg
    # particles per edge
    bppg = int(np.rint(mp.bmap_resolution / (mp.b_length / 3)))
    if bppg < 1:
        raise ValueError("boundary map resolution too fine for the boundary length")
    bx, by = mp.bmap.shape
    mf = _f32(half * _f32(mp.map_factor))     # 1./2. * map_factor
    bp = _f32(bppg)
    xb = []
    nrm = []
    res = mp.bmap_resolution
    for n in range(1, bx + 1):                # DO CONCURRENT (n, m): m fastest,
        for m in range(1, by + 1):            # the compiled reference's order
            if mp.bmap[n - 1, m - 1] != 0:
                continue
            if n + 1 <= bx and mp.bmap[n, m - 1] == 1:
                for k in range(bppg):
                    ax = _f32(n) - mf
                    ay = (_f32(-m + 1) + mf) - (_f32(k) + half) / bp
                    xb.append((mp.bx_mingeom + float(ax) * res, mp.by_maxgeom + float(ay) * res))
                    nrm.append((1.0, 0.0))
            if n - 1 >= 1 and mp.bmap[n - 2, m - 1] == 1:
                for k in range(bppg):
                    ax = _f32(n - 1) - mf
                    ay = (_f32(-m + 1) + mf) - (_f32(k) + half) / bp
                    xb.append((mp.bx_mingeom + float(ax) * res, mp.by_maxgeom + float(ay) * res))
                    nrm.append((-1.0, 0.0))
            if m + 1 <= by and mp.bmap[n - 1, m] == 1:
                for k in range(bppg):
                    ax = (_f32(n - 1) - mf) + (_f32(k) + half) / bp
                    ay = _f32(-m) + mf
                    xb.append((mp.bx_mingeom + float(ax) * res, mp.by_maxgeom + float(ay) * res))
                    nrm.append((0.0, -1.0))
            if m - 1 >= 1 and mp.bmap[n - 1, m - 2] == 1:
                for k in range(bppg):
                    ax = (_f32(n - 1) - mf) + (_f32(k) + half) / bp
                    ay = _f32(-m + 1) + mf
                    xb.append((mp.bx_mingeom + float(ax) * res, mp.by_maxgeom + float(ay) * res))
                    nrm.append((0.0, 1.0))
    xb = np.array(xb, dtype=np.float64).reshape(-1, 2)
    massb = np.full(xb.shape[0], float(np.sum(mass)) / n_p)
    sc = Scenario(x=x, vx=np.zeros((n_p, 2)), mass=mass, rho=rho, hsml=hsml, h=h, A=A,
                  xb=xb, massb=massb, max_hsml=max_hsml, wind=wind, tse=tse,
                  name="reference_default",
                  bnormal=np.array(nrm, dtype=np.float64).reshape(-1, 2))
    sc.dt = vp_dt(hsml) if tse == 2 else 0.0
    return sc
