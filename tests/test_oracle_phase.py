"""Pins of the oracle's ACDI phase field and CSF surface tension
(oracle/mhit_oracle.c, restating multi/main.f90:522-630, 721-780):

- every ACDI term is a flux difference: the box sum of rhs_phi vanishes,
  and phi mass is conserved by full steps;
- a constant velocity advecting phi gives the central difference;
- the diffusive term of a sine is gamma eps (2 cos(k dx) - 2)/dx^2 times it;
- for the planar equilibrium profile 1/2(1 - tanh(x/(2 eps))) the
  sharpening flux cancels most of the diffusive flux (a sign or factor
  slip in either makes them add);
- CSF: zero for uniform phi, aligned with grad phi for a planar profile,
  and the chemical potential of the equilibrium profile is small;
- phi < 0 makes log() return NaN, counted by the diagnostics.
"""
import numpy as np
import pytest

from conftest import relerr
from fields import L, divfree

PHI_CONV, PHI_DIFF, PHI_SHARPEN = 1, 2, 4


def planar(n, eps_r=1.0, width=1.0):
    dx = L / n
    x = np.arange(n) * dx
    prof = 0.5 * (1 - np.tanh((np.abs(x - L / 2) - width) / 2 / (eps_r * dx)))
    return np.ascontiguousarray(np.broadcast_to(prof[None, None, :], (n, n, n)))


@pytest.mark.parametrize("parts", [PHI_CONV, PHI_DIFF, PHI_SHARPEN, 7])
def test_acdi_terms_conserve_mass(parts, oracle_mod):
    n = 16
    rng = np.random.default_rng(parts)
    u, v, w = divfree(n, 10 + parts)
    o = oracle_mod.Oracle(nx=n, phiflag=1)
    for f, a in zip((1, 2, 3), (u, v, w)):
        o.set_field(f, a)
    o.set_field(5, 0.5 + 0.45 * np.sin(3 * rng.standard_normal((n, n, n))))
    r = o.acdi_rhs(parts, 1.7)
    assert abs(r.sum()) < 1e-13 * np.abs(r).sum()
    assert np.abs(r).sum() > 1.0


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_uniform_advection_of_phi(axis, oracle_mod):
    n, kw, U = 32, 4, -0.6
    dx = L / n
    shape = [1, 1, 1]
    shape[2 - axis] = n
    i = np.arange(n)
    phi = np.broadcast_to((0.5 + 0.3 * np.sin(kw * i * dx)).reshape(shape), (n, n, n))
    fields = [np.zeros((n, n, n)) for _ in range(3)]
    fields[axis][:] = U
    o = oracle_mod.Oracle(nx=n, phiflag=1)
    for f, a in zip((1, 2, 3), fields):
        o.set_field(f, a)
    o.set_field(5, phi)
    r = o.acdi_rhs(PHI_CONV, 1.0)
    expect = -U * 0.3 * np.broadcast_to(np.cos(kw * i * dx).reshape(shape), (n, n, n)) \
        * np.sin(kw * dx) / dx
    assert relerr(r, expect) < 1e-12


def test_diffusion_of_a_sine(oracle_mod):
    n, kw, gamma, epsr = 32, 3, 2.5, 1.5
    dx = L / n
    k = np.arange(n)
    s = 0.3 * np.sin(kw * k * dx)[:, None, None]
    o = oracle_mod.Oracle(nx=n, phiflag=1, epsr=epsr)
    o.set_field(5, np.broadcast_to(0.5 + s, (n, n, n)))
    r = o.acdi_rhs(PHI_DIFF, gamma)
    expect = gamma * epsr * dx * (2 * np.cos(kw * dx) - 2) / dx ** 2 * np.broadcast_to(s, (n, n, n))
    assert relerr(r, expect) < 1e-12


def test_sharpening_balances_diffusion_at_equilibrium(oracle_mod):
    n = 64
    o = oracle_mod.Oracle(nx=n, phiflag=1, epsr=2.0)
    o.set_field(5, planar(n, eps_r=2.0))
    diff = o.acdi_rhs(PHI_DIFF, 1.0)
    both = o.acdi_rhs(PHI_DIFF | PHI_SHARPEN, 1.0)
    # second-order residual (0.12 at eps = 2 dx); a flipped sign gives ~2
    assert np.abs(both).max() < 0.2 * np.abs(diff).max()
    # the sharpening flux alone has the opposite sign of the diffusion
    sharp = both - diff
    assert np.sum(sharp * diff) < 0


def test_csf_planar_and_uniform(oracle_mod):
    n = 64
    o = oracle_mod.Oracle(nx=n, phiflag=1, epsr=2.0, sigma=0.7)
    o.set_field(5, planar(n, eps_r=2.0))
    fx, fy, fz = o.csf()
    assert np.abs(fy).max() == 0.0 and np.abs(fz).max() == 0.0
    # chemical potential of the equilibrium profile is small: the force is a
    # small fraction of its bulk-term-only value
    phi = planar(n, eps_r=2.0)
    dx, eps = L / n, 2.0 * L / n
    bulk = 6 * 0.7 * phi * (1 - phi) * (1 - 2 * phi) / eps * \
        0.5 * (np.roll(phi, -1, 2) - np.roll(phi, 1, 2)) / dx
    assert np.abs(fx).max() < 0.05 * np.abs(bulk).max()
    o.set_field(5, np.full((n, n, n), 0.37))
    fx, fy, fz = o.csf()
    assert max(np.abs(fx).max(), np.abs(fy).max(), np.abs(fz).max()) == 0.0


def test_phi_mass_conserved_by_steps(oracle_mod):
    n = 16
    u, v, w = divfree(n, 21)
    o = oracle_mod.Oracle(nx=n, phiflag=1, dt=1e-3, sigma=0.3)
    for f, a in zip((1, 2, 3), (u, v, w)):
        o.set_field(f, a)
    phi = planar(n, eps_r=1.0, width=1.2)
    o.set_field(5, phi)
    m0 = o.diag()["phi_mass"]
    assert o.step(5) == 0
    d = o.diag()
    assert d["phi_mass"] == pytest.approx(m0, rel=1e-13)
    assert not np.allclose(o.get_field(5), phi)


def test_negative_phi_gives_nan(oracle_mod):
    """psi = eps log((min(phi,1)+1e-16)/(1-min(phi,1)+1e-16)) is NaN for
    phi < 0 (main.f90:527-528); the reference does not guard it."""
    n = 8
    o = oracle_mod.Oracle(nx=n, phiflag=1, dt=1e-3, cfl_abort=0)
    phi = np.full((n, n, n), 0.5)
    phi[3, 3, 3] = -0.2
    o.set_field(5, phi)
    o.step(1)
    assert o.diag()["nan_count"] > 0
