"""Pins of the oracle's tracer step (oracle/mhit_oracle.c, restating
multi/particles.f90:78-122, 364-380 and main.f90:477-508):

- a uniform flow is interpolated exactly and moves tracers by dt U
  (Euler, then AB2 with equal velocities), wrapped into [0, 2 pi);
- hand-placed tracers pick the cells of the reference's frame: x index
  floor(x/dx) and the next one (wrapping n -> 0); y and z index
  floor(y/dx) - 1 (the halo-shifted local array) and the next one;
- the trilinear option reproduces any field that is trilinear in the
  staggered coordinates, evaluated by brute force.
"""
import numpy as np
import pytest

from fields import L


def test_uniform_flow_exact(oracle_mod):
    n, dt = 16, 0.05
    U, V, W = 0.9, -1.3, 2.2
    o = oracle_mod.Oracle(nx=n, partflag=1, dt=dt, f1=0.0, f2=0.0, f3=0.0)
    for f, a in zip((1, 2, 3), (U, V, W)):
        o.set_field(f, np.full((n, n, n), a))
    rng = np.random.default_rng(0)
    x0, y0, z0 = (rng.random(500) * L for _ in range(3))
    x0[0], y0[1], z0[2] = L - 1e-3, 1e-4, L - 2e-2   # wrap in both directions
    o.set_tracers(x0, y0, z0)
    assert o.step(1) == 0
    x, y, z, up, vp, wp = o.get_tracers()
    np.testing.assert_allclose(up, U, rtol=1e-14)
    np.testing.assert_allclose(vp, V, rtol=1e-14)
    np.testing.assert_allclose(wp, W, rtol=1e-14)
    # the mean removal (main.f90:977-981) zeroes the uniform flow after the
    # first step; AB2 then moves by dt (1.5*0 - 0.5*U) once, then not at all
    assert o.step(2) == 0
    x, y, z, up, vp, wp = o.get_tracers()
    assert np.abs(up).max() < 1e-12
    for got, start, vel in ((x, x0, U), (y, y0, V), (z, z0, W)):
        expect = np.mod(start + dt * vel - 0.5 * dt * vel, L)
        d = np.abs(got - expect)
        d = np.minimum(d, L - d)
        assert d.max() < 1e-13
        assert got.min() >= 0.0 and got.max() < L


def test_reference_interpolation_cells(oracle_mod):
    n = 8
    dx = L / n
    rng = np.random.default_rng(1)
    u, v, w = (rng.standard_normal((n, n, n)) for _ in range(3))
    o = oracle_mod.Oracle(nx=n, partflag=1)
    for f, a in zip((1, 2, 3), (u, v, w)):
        o.set_field(f, a)
    # (cell index, fraction) per axis; includes the first and the last cell
    cases = [((2, 0.25), (3, 0.5), (4, 0.75)),
             ((7, 0.6), (0, 0.3), (0, 0.9)),
             ((0, 0.1), (7, 0.95), (1, 0.05))]
    pos = np.array([[(c[0] + c[1]) * dx for c in case] for case in cases])
    o.set_tracers(pos[:, 0], pos[:, 1], pos[:, 2])
    up, vp, wp = o.interp()
    for t, ((i, _), (j, _), (k, _)) in enumerate(cases):
        tx, ty, tz = (pos[t, d] / dx - np.floor(pos[t, d] / dx) for d in range(3))
        jm, km = (j - 1) % n, (k - 1) % n          # halo-shifted y/z cell
        # u: along x between cells i and i+1, at y cell j-1, z cell k-1
        assert up[t] == pytest.approx((1 - tx) * u[km, jm, i] + tx * u[km, jm, (i + 1) % n], abs=1e-14)
        # v: along y between cells j-1 and j
        assert vp[t] == pytest.approx((1 - ty) * v[km, jm, i] + ty * v[km, j % n, i], abs=1e-14)
        # w: along z between cells k-1 and k
        assert wp[t] == pytest.approx((1 - tz) * w[km, jm, i] + tz * w[k % n, jm, i], abs=1e-14)


def test_trilinear_option_reproduces_trilinear_field(oracle_mod):
    """u = a + b X + c Y + d Z + e XY + f XZ + g YZ + h XYZ in the
    staggered index coordinates of u, away from the periodic seam."""
    n = 16
    dx = L / n
    coef = np.array([0.3, 0.1, -0.2, 0.05, 0.01, -0.02, 0.03, 0.004])
    f = lambda X, Y, Z: (coef[0] + coef[1] * X + coef[2] * Y + coef[3] * Z + coef[4] * X * Y
                         + coef[5] * X * Z + coef[6] * Y * Z + coef[7] * X * Y * Z)
    idx = np.arange(n, dtype=float)
    Zi, Yi, Xi = np.meshgrid(idx, idx, idx, indexing="ij")
    # storage index (i, j, k) of u sits at ((i - 1/2) dx, (j + 1) dx, (k + 1) dx)
    u = f(Xi - 0.5, Yi + 1.0, Zi + 1.0)
    o = oracle_mod.Oracle(nx=n, partflag=1, interp_mode=1)
    o.set_field(1, u)
    rng = np.random.default_rng(2)
    pos = (3.0 + rng.random((50, 3)) * 8.0) * dx      # away from the seam
    o.set_tracers(pos[:, 0], pos[:, 1], pos[:, 2])
    up, _, _ = o.interp()
    expect = f(pos[:, 0] / dx, pos[:, 1] / dx, pos[:, 2] / dx)
    np.testing.assert_allclose(up, expect, rtol=0, atol=1e-13)
