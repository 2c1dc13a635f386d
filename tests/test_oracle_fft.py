"""Pins of the oracle's FFT and Poisson solve (oracle/mhit_oracle.c,
restating multi/main.f90:853-912): numpy.fft golden vectors, a brute-force
DFT, Parseval, and closed forms of the Poisson solve."""
import os

import numpy as np
import pytest

from conftest import relerr
from fields import L, lap7

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_golden.npz"))


def brute_dft(x, sign):
    n = x.size
    j = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(j, j) / n) @ x


@pytest.mark.parametrize("n", [8, 12, 16, 24, 48, 96])
def test_fft1d_golden(n, oracle_mod):
    x = GOLD[f"fft1d_in_{n}"]
    assert relerr(oracle_mod.fft1d(x, -1), GOLD[f"fft1d_fwd_{n}"]) < 1e-14
    assert relerr(oracle_mod.fft1d(x, +1), GOLD[f"fft1d_inv_{n}"]) < 1e-14


@pytest.mark.parametrize("n", [2, 3, 5, 6, 8, 9, 12, 16])
def test_fft1d_bruteforce(n, oracle_mod):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for sign in (-1, 1):
        assert relerr(oracle_mod.fft1d(x, sign), brute_dft(x, sign)) < 1e-14


@pytest.mark.parametrize("n", [16, 96])
def test_fft1d_f32(n, oracle_mod):
    x = GOLD[f"fft1d_in_{n}"]
    assert relerr(oracle_mod.fft1d_f32(x, -1), GOLD[f"fft1d_fwd_{n}"]) < 1e-5


@pytest.mark.parametrize("n", [8, 12])
def test_fft3d_golden(n, oracle_mod):
    x = GOLD[f"fft3d_in_{n}"]
    assert relerr(oracle_mod.fft3d(x, -1), GOLD[f"fft3d_fwd_{n}"]) < 1e-14


def test_parseval_and_roundtrip(oracle_mod):
    rng = np.random.default_rng(5)
    x = rng.standard_normal((12, 12, 12)) + 1j * rng.standard_normal((12, 12, 12))
    X = oracle_mod.fft3d(x, -1)
    assert np.sum(np.abs(X) ** 2) == pytest.approx(x.size * np.sum(np.abs(x) ** 2), rel=1e-13)
    assert relerr(oracle_mod.fft3d(X, +1) / x.size, x) < 1e-14


@pytest.mark.parametrize("prec,tol", [(0, 1e-14), (1, 1e-6)])
def test_poisson_single_mode(prec, tol, oracle_mod):
    """rhs = cos(m.x + phase) => p = -rhs/|m|^2 (continuous k^2)."""
    n = 24
    xs = np.arange(n) * (L / n)
    Z, Y, X = np.meshgrid(xs, xs, xs, indexing="ij")
    m = (2, -5, 3)
    rhs = np.cos(m[0] * X + m[1] * Y + m[2] * Z + 0.7)
    p = oracle_mod.Oracle(nx=n, poisson_precision=prec).poisson(rhs)
    assert relerr(p, -rhs / sum(a * a for a in m)) < tol


def test_poisson_zero_mode_and_nyquist(oracle_mod):
    """The mean is pinned to zero (main.f90:895); a Nyquist mode along x
    is divided by (n/2)^2 with kx = -n/2 (main.f90:238-240)."""
    n = 16
    i = np.arange(n)
    rhs = np.broadcast_to(((-1.0) ** i)[None, None, :], (n, n, n)) + 3.0
    p = oracle_mod.Oracle(nx=n, poisson_precision=0).poisson(rhs)
    assert relerr(p, -(rhs - 3.0) / (n / 2) ** 2) < 1e-14


@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 2e-6)])
def test_poisson_modified_k2_inverts_7pt_laplacian(prec, tol, oracle_mod):
    """With k'^2 = sum (2 sin(k dx/2)/dx)^2 the solve inverts the discrete
    7-point Laplacian exactly (independent check with numpy.roll)."""
    n = 16
    rng = np.random.default_rng(9)
    rhs = rng.standard_normal((n, n, n))
    rhs -= rhs.mean()
    p = oracle_mod.Oracle(nx=n, poisson_precision=prec, k2_mode=1).poisson(rhs)
    assert relerr(lap7(p, L / n), rhs) < tol
