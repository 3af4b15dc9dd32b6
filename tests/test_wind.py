"""The gridded-wind sampler of the oracle, pinned by what the mathematics
fixes rather than by a second copy of its formula.

The reference's wind is a uniform literal (``ext_force/wind_forcing.f90:
36-37``); the larger workloads carry a static field on a regular grid that the
forcing term of ``wind_forcing.f90:41-49`` samples bilinearly at each
particle, clamped to the grid's edge outside it.  A bilinear sampler is fixed
by four facts, each checked here on a grid with ``nx != ny`` and
``dx != dy`` so that a swapped axis cannot pass:

* it returns the node values exactly at the nodes;
* it reproduces every bilinear field ``a + b x + c y + d x y`` exactly;
* on ``f = x**2`` its error at ``x`` in cell ``[x_i, x_i+1]`` is exactly
  ``(x - x_i) (x_i+1 - x)`` -- which pins the cell it picks (a neighbouring
  cell extrapolates, with the opposite sign);
* outside the grid it returns the value at the nearest edge point.

The synthetic fields themselves (``scenarios.rankine_grid``,
``scenarios.grf_wind``) are checked against their definitions: the Rankine
vortex's analytic speed profile within the interpolation bound, and the
Gaussian random field's standard deviation and Gaussian covariance.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import capi
from simp_amd import scenarios as S

NX, NY = 13, 8
X0, Y0, DX, DY = -3000.0, 1500.0, 2500.0, 4000.0


def _grid(fu, fv):
    gx = X0 + DX * np.arange(NX)
    gy = Y0 + DY * np.arange(NY)
    X, Y = np.meshgrid(gx, gy, indexing="xy")      # (ny, nx), row-major by y
    return S.WindGrid(X0, Y0, DX, DY, np.ascontiguousarray(fu(X, Y)),
                      np.ascontiguousarray(fv(X, Y)))


def _scenario(grid):
    sc = S.box(nside=4, spacing=2000.0, seed=1)
    sc.wind_grid = grid
    return sc


def _inside(rng, k):
    return np.stack([rng.uniform(X0, X0 + DX * (NX - 1), k),
                     rng.uniform(Y0, Y0 + DY * (NY - 1), k)], axis=1)


def test_node_values_are_returned_exactly():
    rng = np.random.default_rng(0)
    U = rng.normal(size=(NY, NX))
    V = rng.normal(size=(NY, NX))
    sc = _scenario(S.WindGrid(X0, Y0, DX, DY, U, V))
    j, i = np.meshgrid(np.arange(NY), np.arange(NX), indexing="ij")
    pts = np.stack([X0 + DX * i.ravel(), Y0 + DY * j.ravel()], axis=1)
    uv = capi.Oracle().wind_at(sc, pts)
    assert np.array_equal(uv[:, 0], U.ravel())
    assert np.array_equal(uv[:, 1], V.ravel())


def test_bilinear_fields_are_reproduced():
    a, b, c, d = 3.0, 2e-4, -5e-4, 3e-9
    fu = lambda X, Y: a + b * X + c * Y + d * X * Y          # noqa: E731
    fv = lambda X, Y: -1.0 + 7e-5 * X - 2e-5 * Y             # noqa: E731
    sc = _scenario(_grid(fu, fv))
    pts = _inside(np.random.default_rng(1), 2000)
    uv = capi.Oracle().wind_at(sc, pts)
    x, y = pts[:, 0], pts[:, 1]
    np.testing.assert_allclose(uv[:, 0], fu(x, y), rtol=0, atol=1e-12)
    np.testing.assert_allclose(uv[:, 1], fv(x, y), rtol=0, atol=1e-12)


@pytest.mark.parametrize("axis", [0, 1])
def test_quadratic_error_pins_the_cell(axis):
    """f = s**2 along one axis: the bilinear sample exceeds s**2 by exactly
    (s - s_i)(s_i+1 - s) >= 0 inside the cell; another cell would give a
    negative excess."""
    if axis == 0:
        sc = _scenario(_grid(lambda X, Y: (X / 1e3) ** 2, lambda X, Y: 0.0 * X))
        h, o = DX / 1e3, X0 / 1e3
    else:
        sc = _scenario(_grid(lambda X, Y: 0.0 * X, lambda X, Y: (Y / 1e3) ** 2))
        h, o = DY / 1e3, Y0 / 1e3
    pts = _inside(np.random.default_rng(2 + axis), 3000)
    uv = capi.Oracle().wind_at(sc, pts)
    s = pts[:, axis] / 1e3
    got = uv[:, axis]
    i = np.floor((s - o) / h)
    lo, hi = o + h * i, o + h * (i + 1)
    want = s * s + (s - lo) * (hi - s)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-9)
    assert np.all(got - s * s >= -1e-9)
    # the other component of the same field is identically zero
    assert np.all(uv[:, 1 - axis] == 0.0)


def test_points_outside_are_clamped_to_the_edge():
    rng = np.random.default_rng(3)
    U = rng.normal(size=(NY, NX))
    V = rng.normal(size=(NY, NX))
    sc = _scenario(S.WindGrid(X0, Y0, DX, DY, U, V))
    o = capi.Oracle()
    xe, ye = X0 + DX * (NX - 1), Y0 + DY * (NY - 1)
    inner = _inside(rng, 400)
    for sx, sy in ((-1, 0), (1, 0), (0, -1), (0, 1), (-1, -1), (1, 1), (-1, 1), (1, -1)):
        out = inner.copy()
        if sx < 0:
            out[:, 0] = X0 - rng.uniform(1.0, 5e4, len(out))
        if sx > 0:
            out[:, 0] = xe + rng.uniform(1.0, 5e4, len(out))
        if sy < 0:
            out[:, 1] = Y0 - rng.uniform(1.0, 5e4, len(out))
        if sy > 0:
            out[:, 1] = ye + rng.uniform(1.0, 5e4, len(out))
        edge = np.stack([np.clip(out[:, 0], X0, xe), np.clip(out[:, 1], Y0, ye)], axis=1)
        np.testing.assert_array_equal(o.wind_at(sc, out), o.wind_at(sc, edge))


def test_uniform_mode_ignores_position():
    sc = S.box(nside=4, spacing=2000.0, seed=1)
    sc.wind = (4.5, -2.25)
    uv = capi.Oracle().wind_at(sc, _inside(np.random.default_rng(4), 50))
    assert np.all(uv[:, 0] == 4.5) and np.all(uv[:, 1] == -2.25)


def test_rankine_vortex_matches_its_definition():
    """The vortex workload's field sampled at its particles: counter-clockwise
    around the box centre, speed V r / R inside R and V R / r outside, within
    the bilinear interpolation bound (h^2 / 8 max|f''| per axis, taken
    generously near the core where the profile has a kink)."""
    sc = S.vortex(nside=60, spacing=1000.0, seed=2)       # 60 km box
    L = 60e3
    R = 60000.0 * (L / 200000.0)
    g = sc.wind_grid
    uv = capi.Oracle().wind_at(sc, sc.x)
    rx, ry = sc.x[:, 0] - L / 2, sc.x[:, 1] - L / 2
    r = np.hypot(rx, ry)
    spd = np.where(r <= R, 15.0 * r / R, 15.0 * R / r)
    want = np.stack([-spd * ry / r, spd * rx / r], axis=1)
    err = np.abs(uv - want).max(axis=1)
    # away from the core's kink the field is smooth: |f''| <= ~3 V / R^2
    smooth = np.abs(r - R) > 1.5 * g.dx
    bound = 3.0 * 15.0 / R ** 2 * (g.dx ** 2 + g.dy ** 2) / 8 * 4
    assert np.all(err[smooth] <= bound), err[smooth].max()
    # at the kink the error is still a fraction of the grid's speed change
    assert np.all(err <= 15.0 * g.dx / R)
    # rotation sense: counter-clockwise (positive angular momentum)
    assert np.all(rx * uv[:, 1] - ry * uv[:, 0] > 0)


def test_gaussian_random_field_statistics():
    """scenarios.grf_wind: zero mean and the stated standard deviation per
    component (normalised exactly), and a covariance that follows
    exp(-r^2 / (2 l^2)) -- checked on the ensemble of many fields at lags
    0.5 l, l and 2 l, with a tolerance several times the sampling error."""
    L, std, corr, gdx = 400e3, 8.0, 50e3, 5e3
    rng = np.random.default_rng(11)
    lags = {5: 0.5, 10: 1.0, 20: 2.0}        # in grid cells -> r / l
    acc = {k: [] for k in lags}
    for _ in range(12):
        g = S.grf_wind(L, std, corr, gdx, rng)
        for f in (g.u, g.v):
            assert abs(float(f.mean())) < 1e-9
            assert abs(float(f.std()) - std) < 1e-9
            for k in lags:
                c = 0.5 * (np.mean(f[:, k:] * f[:, :-k]) + np.mean(f[k:, :] * f[:-k, :]))
                acc[k].append(c / std ** 2)
    for k, rl in lags.items():
        got = float(np.mean(acc[k]))
        want = math.exp(-0.5 * rl * rl)
        assert abs(got - want) < 0.08, (k, got, want)


def test_ice_edge_wind_sampled_on_its_grid():
    """The headline workload's field: every particle lies inside the grid,
    and the sample there is a convex combination of the four surrounding
    node values (bounded by their min and max)."""
    sc = S.ice_edge(nside=100, spacing=1000.0, seed=3)
    g = sc.wind_grid
    uv = capi.Oracle().wind_at(sc, sc.x)
    ny, nx = g.u.shape
    fx = (sc.x[:, 0] - g.x0) / g.dx
    fy = (sc.x[:, 1] - g.y0) / g.dy
    assert fx.min() >= 0 and fy.min() >= 0 and fx.max() <= nx - 1 and fy.max() <= ny - 1
    i = np.minimum(np.floor(fx).astype(int), nx - 2)
    j = np.minimum(np.floor(fy).astype(int), ny - 2)
    for c, F in ((0, g.u), (1, g.v)):
        corners = np.stack([F[j, i], F[j, i + 1], F[j + 1, i], F[j + 1, i + 1]], axis=1)
        assert np.all(uv[:, c] >= corners.min(axis=1) - 1e-12)
        assert np.all(uv[:, c] <= corners.max(axis=1) + 1e-12)
