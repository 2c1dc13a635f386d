"""Pins of the oracle's Navier-Stokes part (oracle/mhit_oracle.c, restating
multi/main.f90:658-717, 783-802, 833-1003) by closed forms and invariants:

- the divergence-form convective term of the staggered grid conserves
  discrete kinetic energy for a discretely divergence-free field;
- a constant velocity advecting a sine gives the central difference
  -U (f(+1) - f(-1))/(2 dx) = -U cos(kx) sin(k dx)/dx for each of the six
  cross terms h12, h13, h21, h23, h31, h32;
- the viscous term of a sine is mu (2 cos(k dx) - 2)/dx^2 times the sine;
- the ABC forcing table in the time-stepper frame (y, z of interior index
  j are (j+1) dx) and zero box sum;
- Euler then AB2 reproduce the exact recurrence of a decaying mode;
- the projection: D.u_new = (1 - k'^2/k^2) D.u* mode by mode with the
  continuous k^2, and D.u_new = 0 with the modified k'^2;
- signed gumax, cou, and the mean removal.
"""
import numpy as np
import pytest

from conftest import relerr
from fields import L, divfree, fdiv

NS_CONV, NS_VISC, NS_FORCE = 1, 2, 4


def oracle(oracle_mod, n, **kw):
    return oracle_mod.Oracle(nx=n, **kw)


def test_convective_energy_conservation(oracle_mod):
    n = 16
    u, v, w = divfree(n, 1)
    assert np.abs(fdiv(u, v, w)).max() < 1e-13
    o = oracle(oracle_mod, n)
    for f, a in zip((1, 2, 3), (u, v, w)):
        o.set_field(f, a)
    cu, cv, cw = o.ns_rhs(NS_CONV)
    transfer = np.sum(u * cu + v * cv + w * cw)
    scale = np.sum(np.abs(u * cu) + np.abs(v * cv) + np.abs(w * cw))
    assert abs(transfer) < 1e-13 * scale
    # and the term is not trivially zero
    assert scale > 1.0


@pytest.mark.parametrize("carrier,target,axis", [
    (0, 1, 0),  # u=U advects v along x (h21)
    (1, 0, 1),  # v=V advects u along y (h12)
    (2, 0, 2),  # w=W advects u along z (h13)
    (0, 2, 0),  # u=U advects w along x (h31)
    (2, 1, 2),  # w=W advects v along z (h23)
    (1, 2, 1),  # v=V advects w along y (h32)
])
def test_uniform_advection_central_difference(carrier, target, axis, oracle_mod):
    n, kw, U = 32, 3, 0.7
    dx = L / n
    idx = np.arange(n)
    shape = [1, 1, 1]
    shape[2 - axis] = n                     # array axis of physical axis
    s = np.sin(kw * idx * dx).reshape(shape)
    c = np.cos(kw * idx * dx).reshape(shape)
    fields = [np.zeros((n, n, n)) for _ in range(3)]
    fields[carrier][:] = U
    fields[target][:] = np.broadcast_to(s, (n, n, n))
    o = oracle(oracle_mod, n)
    for f, a in zip((1, 2, 3), fields):
        o.set_field(f, a)
    rhs = o.ns_rhs(NS_CONV)
    expect = -U * np.broadcast_to(c, (n, n, n)) * np.sin(kw * dx) / dx
    assert relerr(rhs[target], expect) < 1e-12
    assert np.abs(rhs[carrier]).max() < 1e-12


@pytest.mark.parametrize("comp,axis", [(0, 0), (0, 1), (1, 2), (2, 1), (2, 0)])
def test_viscous_term_of_a_sine(comp, axis, oracle_mod):
    n, kw, mu, rho = 32, 5, 0.013, 1.7
    dx = L / n
    shape = [1, 1, 1]
    shape[2 - axis] = n
    s = np.broadcast_to(np.sin(kw * np.arange(n) * dx + 0.2).reshape(shape), (n, n, n))
    fields = [np.zeros((n, n, n)) for _ in range(3)]
    fields[comp][:] = s
    o = oracle(oracle_mod, n, mu=mu, rho=rho)
    for f, a in zip((1, 2, 3), fields):
        o.set_field(f, a)
    rhs = o.ns_rhs(NS_VISC)
    lam = mu * (2 * np.cos(kw * dx) - 2) / dx ** 2 / rho
    assert relerr(rhs[comp], lam * s) < 1e-12


def test_abc_forcing_frame(oracle_mod):
    """rhs_u = C sin(k0 z + dx/2) + B cos(k0 y + dx/2),
    rhs_v = A sin(k0 x + dx/2) + C cos(k0 z + dx/2),
    rhs_w = B sin(k0 y + dx/2) + A cos(k0 x + dx/2), with x = i dx and
    y, z = (j+1) dx, (k+1) dx for interior index (i, j, k) (main.f90:709-714
    with jg = j of the halo-shifted array)."""
    n, k0 = 16, 2.0
    A, B, C = 1.0, 2.0, 3.0
    dx = L / n
    o = oracle(oracle_mod, n, f1=A, f2=B, f3=C, k0=k0)
    ru, rv, rw = o.ns_rhs(NS_FORCE)
    i = np.arange(n)
    x = (i * dx)[None, None, :]
    y = ((i + 1) * dx)[None, :, None]
    z = ((i + 1) * dx)[:, None, None]
    sh = lambda a: np.broadcast_to(a, (n, n, n))
    assert relerr(ru, sh(C * np.sin(k0 * z + dx / 2) + B * np.cos(k0 * y + dx / 2))) < 1e-13
    assert relerr(rv, sh(A * np.sin(k0 * x + dx / 2) + C * np.cos(k0 * z + dx / 2))) < 1e-13
    assert relerr(rw, sh(B * np.sin(k0 * y + dx / 2) + A * np.cos(k0 * x + dx / 2))) < 1e-13
    for r in (ru, rv, rw):
        assert abs(r.sum()) < 1e-10


def test_euler_then_ab2_recurrence(oracle_mod):
    """u = a sin(m y), v = w = 0, no forcing: convection, divergence and
    pressure vanish, so u follows du/dt = lam u with Euler for step 1 and
    AB2 afterwards (main.f90:440-441, 783-802)."""
    n, m, mu, dt = 16, 3, 0.05, 1e-3
    dx = L / n
    j = np.arange(n)
    prof = np.sin(m * j * dx)[None, :, None]
    u0 = np.ascontiguousarray(np.broadcast_to(0.8 * prof, (n, n, n)))
    o = oracle(oracle_mod, n, mu=mu, dt=dt, f1=0.0, f2=0.0, f3=0.0, poisson_precision=0)
    o.set_field(1, u0)
    lam = mu * (2 * np.cos(m * dx) - 2) / dx ** 2
    a = [0.8]
    r = []
    for step in range(6):
        r.append(lam * a[-1])
        if step == 0:
            a.append(a[-1] + dt * r[-1])
        else:
            a.append(a[-1] + dt * (1.5 * r[-1] - 0.5 * r[-2]))
    assert o.step(6) == 0
    u = o.get_field(1)
    assert relerr(u, a[-1] / 0.8 * u0) < 1e-13
    d = o.diag()
    assert d["alpha"] == 1.5 and d["beta"] == 0.5


@pytest.mark.parametrize("prec", [0, 1])
def test_projection_closed_form(prec, oracle_mod):
    """p = solve(rho/(dt dx) D.u*), u = u* - dt/rho grad_b p  =>
    F[D.u] = (1 - k'^2/k^2) F[D.u*] with k'^2 = sum (2 sin(k dx/2)/dx)^2."""
    n, dt, rho = 16, 1e-3, 1.3
    dx = L / n
    rng = np.random.default_rng(4)
    u, v, w = (rng.standard_normal((n, n, n)) for _ in range(3))
    o = oracle(oracle_mod, n, dt=dt, rho=rho, poisson_precision=prec)
    for f, a in zip((1, 2, 3), (u, v, w)):
        o.set_field(f, a)
    rhsp = o.divergence()
    assert relerr(rhsp, rho / (dt * dx) * fdiv(u, v, w)) < 1e-13
    p = o.poisson(rhsp)
    g = lambda ax: (p - np.roll(p, 1, ax)) / dx
    un, vn, wn = u - dt / rho * g(2), v - dt / rho * g(1), w - dt / rho * g(0)
    k = np.fft.fftfreq(n, 1.0 / n)
    KZ, KY, KX = np.meshgrid(k, k, k, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    kp2 = sum((2 * np.sin(K * dx / 2) / dx) ** 2 for K in (KX, KY, KZ))
    k2[0, 0, 0] = 1.0
    factor = 1 - kp2 / k2
    factor[0, 0, 0] = 1.0
    lhs = np.fft.fftn(fdiv(un, vn, wn))
    rhs = factor * np.fft.fftn(fdiv(u, v, w))
    assert relerr(lhs, rhs) < (1e-12 if prec == 0 else 1e-5)


def test_full_step_modified_k2_is_divergence_free(oracle_mod):
    n = 16
    rng = np.random.default_rng(8)
    o = oracle(oracle_mod, n, k2_mode=1, poisson_precision=0)
    for f in (1, 2, 3):
        o.set_field(f, rng.standard_normal((n, n, n)))
    assert o.step(2) == 0
    u, v, w = (o.get_field(f) for f in (1, 2, 3))
    assert np.abs(fdiv(u, v, w)).max() / (L / n) < 1e-10
    assert abs(u.mean()) < 1e-15 and abs(v.mean()) < 1e-15 and abs(w.mean()) < 1e-15


def test_gumax_cou_and_mean(oracle_mod):
    """gumax is the signed max over u, v, w after the mean removal
    (main.f90:977-1003) and cou = gumax dt / dx."""
    n, dt = 16, 2e-3
    u, v, w = divfree(n, 2)
    o = oracle(oracle_mod, n, dt=dt)
    for f, a in zip((1, 2, 3), (u + 0.3, v - 0.2, w)):
        o.set_field(f, a)
    assert o.step(1) == 0
    un, vn, wn = (o.get_field(f) for f in (1, 2, 3))
    d = o.diag()
    assert d["gumax"] == max(un.max(), vn.max(), wn.max())
    assert d["cou"] == pytest.approx(d["gumax"] * dt / (L / n), rel=1e-15)
    # every right-hand-side term and the pressure gradient sum to zero over
    # the periodic box, so the removed mean is the initial one
    assert d["umean"] == pytest.approx((u + 0.3).mean(), abs=1e-14)
    assert d["vmean"] == pytest.approx((v - 0.2).mean(), abs=1e-14)
    for a in (un, vn, wn):
        assert abs(a.mean()) < 1e-15


def test_cfl_stop(oracle_mod):
    n = 8
    o = oracle(oracle_mod, n, dt=1.0)
    o.set_field(1, np.full((n, n, n), 10.0) + divfree(n, 3)[0])
    assert o.step(3) == 3   # MHIT_ERR_CFL after the first step
    assert o.diag()["step"] == 1


def test_taylor_green_diagnostics(oracle_mod):
    """TGV of main.f90:324-337: KE = 1/8 on the grid; dissipation with
    central differences tends to 3 nu/4 (van Rees et al. t=0 value)."""
    n, mu = 64, 6.25e-4
    dx = L / n
    xe = np.arange(n + 1) * dx
    x = xe[:n][None, None, :]
    y = xe[1:][None, :, None]
    z = xe[1:][:, None, None]
    u = np.sin(x - dx / 2) * np.cos(y) * np.cos(z)
    v = -np.cos(x) * np.sin(y - dx / 2) * np.cos(z)
    o = oracle(oracle_mod, n, mu=mu)
    o.set_field(1, np.broadcast_to(u, (n, n, n)).copy())
    o.set_field(2, np.broadcast_to(v, (n, n, n)).copy())
    d = o.diag()
    assert d["ke"] == pytest.approx(0.125, rel=1e-13)
    assert d["dissipation"] == pytest.approx(3 * mu / 4, rel=3e-3)
