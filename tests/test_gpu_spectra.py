"""Kinetic-energy shell spectrum on the GPU against numpy.fft (a library
routine) on the downloaded velocity, and against the kinetic energy of the
diagnostics (Parseval)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _numpy_spectrum(u, v, w):
    n = u.shape[0]
    N = n ** 3
    k = np.fft.fftfreq(n, 1.0 / n)
    kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")
    km = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    e = sum(np.abs(np.fft.fftn(a)) ** 2 for a in (u, v, w)) * 0.5 / N ** 2
    shells = np.floor(km + 0.5).astype(int)
    nb = int(np.floor(np.sqrt(3.0) * (n // 2) + 0.5)) + 1
    return np.bincount(shells.ravel(), weights=e.ravel(), minlength=nb)


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
@pytest.mark.parametrize("n", [16, 48, 64])
def test_spectrum_matches_numpy(n, prec):
    from mhit36_mixedp_amd import Solver
    print(f"[spectrum] n={n} prec={prec}", flush=True)
    with Solver(nx=n, poisson_precision=prec, dt=5e-4) as s:
        s.init_random_velocity(11, kmax=6.0, urms=1.0)
        s.step(3)                               # a non-zero deferred mean
        u, v, w = s.velocity()
        E = s.energy_spectrum()
        ref = _numpy_spectrum(u, v, w)
        tol = 1e-12 if prec == 0 else 1e-6
        assert E.shape == ref.shape
        assert np.max(np.abs(E - ref)) < tol * ref.sum()
        assert abs(E.sum() - s.diag()["ke"]) < tol * E.sum()


def test_single_mode_shell():
    """u = cos(3x + 4y): all energy 1/4 in the |k| = 5 shell."""
    from mhit36_mixedp_amd import Solver
    print("[spectrum] single mode", flush=True)
    n = 32
    x = np.arange(n) * (2 * np.pi / n)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    with Solver(nx=n, poisson_precision=0) as s:
        s.set_velocity(np.cos(3 * X + 4 * Y), np.zeros_like(X), np.zeros_like(X))
        E = s.energy_spectrum()
    assert E[5] == pytest.approx(0.25, rel=1e-13)
    assert np.sum(np.abs(np.delete(E, 5))) < 1e-15
