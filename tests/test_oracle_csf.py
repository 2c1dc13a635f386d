"""Pins of the oracle's CSF coupling (oracle/mhit_oracle.c `csf` and the
face average of `one_step`, restating multi/main.f90:721-743, 760-780) by
what the mathematics fixes, not by restating the formula:

- Young-Laplace: a static drop of radius R, started from rest with no
  forcing, is projected in one step to a pressure field whose jump across
  the interface is 2 sigma / R.  For the tanh profile of width eps the
  force 6 sigma mu_c grad(phi), mu_c = phi(1-phi)(1-2phi)/eps - eps lap(phi),
  integrates to exactly that jump (the planar part of mu_c vanishes and the
  curvature part gives 12 sigma eps/R int phi'^2 dr = 2 sigma/R).  This
  fixes the 6 sigma amplitude and the 1/2 factors of the central gradient
  and of the face average.
- Mirror antisymmetry: for a drop centred on a cell centre c, the face
  velocity u(i) (at x_{i-1/2}) after one step from rest satisfies
  u(c+m) = -u(c+1-m) along its own axis, and likewise v in y and w in z.
  Placing the cell-centred force on the faces with 1/2(f(i-1) + f(i)) is
  what makes this hold; a half-cell slip (f(i+1)) breaks it at O(1).
"""
import numpy as np
import pytest

from fields import L


def drop_field(n, radius_cells, epsr, centre):
    """phi = 1/2(1 - tanh((r - R)/(2 eps))) about the cell index `centre`,
    r the periodic minimum-image distance."""
    dx = L / n
    d = (np.arange(n) - centre + n // 2) % n - n // 2
    r = np.sqrt(d[None, None, :] ** 2 + d[None, :, None] ** 2 + d[:, None, None] ** 2) * dx
    return 0.5 * (1.0 - np.tanh((r - radius_cells * dx) / 2.0 / (epsr * dx))), r


def step_from_rest(oracle_mod, n, phi, **kw):
    o = oracle_mod.Oracle(nx=n, phiflag=1, partflag=0, poisson_precision=0,
                          f1=0.0, f2=0.0, f3=0.0, **kw)
    for f in (1, 2, 3):
        o.set_field(f, np.zeros((n, n, n)))
    o.set_field(5, phi)
    assert o.step(1) == 0
    return o


@pytest.mark.parametrize("sigma,radius_cells,epsr,k2,rho", [
    (1.0, 10, 1.0, 0, 1.0),
    (0.35, 10, 1.0, 1, 1.0),
    (1.0, 8, 1.5, 0, 1.0),
    (2.5, 12, 1.0, 0, 2.0),
])
def test_young_laplace_pressure_jump(oracle_mod, sigma, radius_cells, epsr, k2, rho):
    n = 64
    dx = L / n
    phi, r = drop_field(n, radius_cells, epsr, n // 2)
    # the jump is independent of rho: u* = dt f/rho, rhs_p = rho/dt div u*
    o = step_from_rest(oracle_mod, n, phi, sigma=sigma, epsr=epsr, k2_mode=k2,
                       dt=1e-4, mu=0.006, rho=rho)
    p = o.get_field(4)
    R, eps = radius_cells * dx, epsr * dx
    inside, outside = p[r < R - 4 * eps], p[r > R + 6 * eps]
    jump = inside.mean() - outside.mean()
    expect = 2.0 * sigma / R
    # measured 0.97-1.01 of 2 sigma/R (second-order interface error at
    # eps = dx); a factor slip moves it by 2x
    assert jump == pytest.approx(expect, rel=0.05)
    # the pressure is flat away from the interface on both sides
    assert inside.std() < 0.01 * expect and outside.std() < 0.01 * expect


@pytest.mark.parametrize("n,centre,radius_cells", [(32, 15, 6), (24, 9, 5)])
def test_face_force_mirror_antisymmetry(oracle_mod, n, centre, radius_cells):
    phi, _ = drop_field(n, radius_cells, 1.0, centre)
    o = step_from_rest(oracle_mod, n, phi, sigma=1.0, dt=1e-3, mu=0.006)
    u, v, w = (o.get_field(f) for f in (1, 2, 3))
    i = np.arange(n)
    face = (2 * centre + 1 - i) % n     # x_{c+m-1/2} <-> x_{c-m+1/2}
    cell = (2 * centre - i) % n         # cell-centred mirror
    scale = max(np.abs(a).max() for a in (u, v, w))
    assert scale > 1e-6                 # the surface tension moved the fluid
    # each component is antisymmetric about the centre along its own axis ...
    assert np.abs(u + u[:, :, face]).max() < 1e-10 * scale
    assert np.abs(v + v[:, face, :]).max() < 1e-10 * scale
    assert np.abs(w + w[face, :, :]).max() < 1e-10 * scale
    # ... and symmetric along the other two
    assert np.abs(u - u[:, cell, :]).max() < 1e-10 * scale
    assert np.abs(u - u[cell, :, :]).max() < 1e-10 * scale
    assert np.abs(v - v[:, :, cell]).max() < 1e-10 * scale
    assert np.abs(w - w[:, :, cell]).max() < 1e-10 * scale
