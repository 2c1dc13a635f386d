"""GPU FFT / Poisson parity (HIP path through the C-ABI).

Pins: numpy.fft (library routine) for the 3-D R2C/C2R transforms, the
closed-form single-mode Poisson solution, and the CPU oracle's Poisson
solve (multi/main.f90:853-912).  Tolerances: FP64 path 1e-12 relative to
the field's max; FP32 path 2e-6 (transform) / 1e-5 (solve) relative."""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu

SIZES = [8, 12, 16, 24, 32, 48, 64, 96, 128]


def _solver(n, prec, **kw):
    from mhit36_mixedp_amd import Solver
    return Solver(nx=n, poisson_precision=prec, **kw)


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
@pytest.mark.parametrize("n", SIZES)
def test_r2c_matches_numpy(n, prec):
    print(f"[fft r2c] n={n} prec={prec}", flush=True)
    rng = np.random.default_rng(100 + n)
    x = rng.standard_normal((n, n, n))
    with _solver(n, prec) as s:
        X = s.fft_r2c(x)
    ref = np.fft.rfftn(x)
    tol = 1e-12 if prec == 0 else 2e-6
    assert relerr(X, ref) < tol, relerr(X, ref)


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
@pytest.mark.parametrize("n", SIZES)
def test_c2r_matches_numpy(n, prec):
    print(f"[fft c2r] n={n} prec={prec}", flush=True)
    rng = np.random.default_rng(200 + n)
    x = rng.standard_normal((n, n, n))
    X = np.fft.rfftn(x)
    with _solver(n, prec) as s:
        y = s.fft_c2r(X)
    ref = np.fft.irfftn(X, s=(n, n, n), axes=(0, 1, 2)) * n ** 3
    tol = 1e-12 if prec == 0 else 2e-6
    assert relerr(y, ref) < tol, relerr(y, ref)


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
@pytest.mark.parametrize("n", [16, 32, 96])
def test_poisson_single_mode(n, prec):
    """rhs = cos(m.x + 0.3) => p = -rhs/|m|^2 for the continuous k^2."""
    print(f"[poisson mode] n={n} prec={prec}", flush=True)
    L = 2 * np.pi
    xs = np.arange(n) * (L / n)
    Z, Y, X = np.meshgrid(xs, xs, xs, indexing="ij")
    m = (3, -2, 5)
    rhs = np.cos(m[0] * X + m[1] * Y + m[2] * Z + 0.3)
    with _solver(n, prec) as s:
        p = s.poisson_solve(rhs)
    ref = -rhs / sum(a * a for a in m)
    tol = 1e-13 if prec == 0 else 2e-6
    assert relerr(p, ref) < tol


def _modes(n):
    h = n // 2
    return [(3, -2, 5), (h, 1, -3), (-h + 1, h - 1, 7), (0, 0, h), (1, h, 0),
            (n // 4, -(n // 4), h - 3)]


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
@pytest.mark.parametrize("n", [256, 512])
def test_poisson_modes_full_size(n, prec):
    """Six modes at production sizes, including x, y and z Nyquist modes
    and the highest wavenumbers: rhs = sum cos(m.x + ph) gives
    p = -sum cos(m.x + ph)/|m|^2 for the continuous k^2 (main.f90:883-896).
    The tolerance is below the smallest mode's amplitude in p, so a mode
    sent to the wrong wavenumber fails."""
    print(f"[poisson modes] n={n} prec={prec}", flush=True)
    L = 2 * np.pi
    xs = np.arange(n) * (L / n)
    rhs = np.zeros((n, n, n))
    ref = np.zeros((n, n, n))
    amp = []
    for j, m in enumerate(_modes(n)):
        ph = 0.3 + 0.7 * j
        e = (np.exp(1j * m[2] * xs)[:, None, None] * np.exp(1j * m[1] * xs)[None, :, None]
             * (np.exp(1j * (m[0] * xs + ph)))[None, None, :]).real
        k2 = sum(a * a for a in m)
        rhs += e
        ref -= e / k2
        amp.append(1.0 / k2)
        del e
    with _solver(n, prec) as s:
        p = s.poisson_solve(rhs)
    err = np.abs(p - ref).max()
    tol = 1e-13 if prec == 0 else 1e-6
    print(f"   max err {err:.3e}, smallest mode amplitude {min(amp):.3e}", flush=True)
    assert tol < 0.2 * min(amp)
    assert err < tol


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
@pytest.mark.parametrize("k2", [0, 1], ids=["k2", "k2mod"])
@pytest.mark.parametrize("n", [32, 64])
def test_poisson_vs_oracle(n, prec, k2, oracle_mod):
    print(f"[poisson oracle] n={n} prec={prec} k2={k2}", flush=True)
    rng = np.random.default_rng(7 + n)
    rhs = rng.standard_normal((n, n, n))
    rhs -= rhs.mean()
    with _solver(n, prec, k2_mode=k2) as s:
        p = s.poisson_solve(rhs)
    o = oracle_mod.Oracle(nx=n, poisson_precision=prec, k2_mode=k2)
    ref = o.poisson(rhs)
    tol = 1e-12 if prec == 0 else 1e-5
    assert relerr(p, ref) < tol, relerr(p, ref)


@pytest.mark.parametrize("form", [0, 1], ids=["direct", "wavefront"])
@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
@pytest.mark.parametrize("n", [64, 128, 256])
def test_both_transform_forms_match_numpy(n, prec, form):
    """Both forms of the column transforms (Tuning::fft_wave) against
    numpy: forward (r2c: x, y, z passes), inverse (c2r) and the z pass's
    forward + divide + inverse (poisson_solve against the spectral
    solution p^ = -r^/|k|^2 computed with numpy)."""
    print(f"[fft forms] n={n} prec={prec} form={form}", flush=True)
    rng = np.random.default_rng(300 + n + form)
    x = rng.standard_normal((n, n, n))
    tol = 1e-12 if prec == 0 else 2e-6
    with _solver(n, prec) as s:
        s.set_tuning(fft_wave=form)
        assert s.tuning()["fft_wave"] == form
        X = s.fft_r2c(x)
        assert relerr(X, np.fft.rfftn(x)) < tol
        y = s.fft_c2r(np.fft.rfftn(x))
        assert relerr(y, x * n ** 3) < tol
        p = s.poisson_solve(x)
    k = np.fft.fftfreq(n, 1.0 / n)
    kz, ky, kx = np.meshgrid(k, k, k[: n // 2 + 1], indexing="ij")
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    k2[0, 0, 0] = 1.0
    ph = -np.fft.rfftn(x) / k2
    ph[0, 0, 0] = 0.0
    ref = np.fft.irfftn(ph, s=(n, n, n), axes=(0, 1, 2))
    assert relerr(p, ref) < (1e-12 if prec == 0 else 1e-5)


@pytest.mark.parametrize("form", [0, 1], ids=["direct", "wavefront"])
@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
@pytest.mark.parametrize("n", [64, 96, 256])
def test_every_launch_width_matches_numpy(n, prec, form):
    """The alternative FFT launch widths (Tuning fft_ycols / fft_zcols /
    fft_xrows, one at a time) against numpy: forward, inverse and the
    Poisson solve, in both transform forms (eight columns take the
    wavefront form when it is on), x and y transforms in separate
    launches."""
    print(f"[fft widths] n={n} prec={prec} form={form}", flush=True)
    rng = np.random.default_rng(700 + n + form)
    x = rng.standard_normal((n, n, n))
    tol = 1e-12 if prec == 0 else 2e-6
    k = np.fft.fftfreq(n, 1.0 / n)
    kz, ky, kx = np.meshgrid(k, k, k[: n // 2 + 1], indexing="ij")
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    k2[0, 0, 0] = 1.0
    ph = -np.fft.rfftn(x) / k2
    ph[0, 0, 0] = 0.0
    ref = np.fft.irfftn(ph, s=(n, n, n), axes=(0, 1, 2))
    with _solver(n, prec) as s:
        base = s.tuning()
        alts = [dict(fft_ycols=base["fft_ycols"] // 2),
                dict(fft_zcols=base["fft_zcols"] * 2 if prec == 1 else base["fft_zcols"] // 2),
                dict(fft_xrows=base["fft_xrows"] // 2)]
        if (n // 2) % (2 * base["fft_xrows"]) == 0:
            alts.append(dict(fft_xrows=base["fft_xrows"] * 2))
        for alt in alts:
            kw = dict(fft_wave=form, fft_ycols=0, fft_zcols=0, fft_xrows=0, fft_xy=0)
            kw.update(alt)
            s.set_tuning(**kw)
            assert all(s.tuning()[kk] == v for kk, v in alt.items())
            assert relerr(s.fft_r2c(x), np.fft.rfftn(x)) < tol, alt
            assert relerr(s.fft_c2r(np.fft.rfftn(x)), x * n ** 3) < tol, alt
            assert relerr(s.poisson_solve(x), ref) < (1e-12 if prec == 0 else 1e-5), alt


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
@pytest.mark.parametrize("n", [32, 48, 64, 96, 128])
def test_xy_in_one_launch_and_apart_match_numpy(n, prec):
    """The x and y transforms of a plane in one launch (fft_xy = 1, the
    default where compiled) and in two: forward, inverse and the Poisson
    solve against numpy, and the two forms against each other at the
    transform tolerance.  The one-launch form packs the kx = 0 and
    kx = n/2 columns into one complex column and separates them after the
    y transform, so a wrong pack or split fails the Nyquist plane of the
    forward transform and the real rows of the inverse."""
    print(f"[fft xy] n={n} prec={prec}", flush=True)
    rng = np.random.default_rng(900 + n + prec)
    x = rng.standard_normal((n, n, n))
    tol = 1e-12 if prec == 0 else 2e-6
    X = np.fft.rfftn(x)
    k = np.fft.fftfreq(n, 1.0 / n)
    kz, ky, kx = np.meshgrid(k, k, k[: n // 2 + 1], indexing="ij")
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    k2[0, 0, 0] = 1.0
    ph = -X / k2
    ph[0, 0, 0] = 0.0
    ref = np.fft.irfftn(ph, s=(n, n, n), axes=(0, 1, 2))
    out = {}
    with _solver(n, prec) as s:
        for v in (1, 0):
            s.set_tuning(fft_xy=v)
            assert s.tuning()["fft_xy"] == (v if (prec == 1 or n <= 96) else 0)
            F = s.fft_r2c(x)
            assert relerr(F, X) < tol, v
            assert relerr(F[:, :, n // 2], X[:, :, n // 2]) < tol, v   # the Nyquist plane
            y = s.fft_c2r(X)
            assert relerr(y, x * n ** 3) < tol, v
            p = s.poisson_solve(x)
            assert relerr(p, ref) < (1e-12 if prec == 0 else 1e-5), v
            out[v] = (F, y, p)
    for a, b in zip(out[1], out[0]):
        assert relerr(a, b) < tol


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
@pytest.mark.parametrize("n", [128, 256])
def test_x_radix_plans_match_numpy(n, prec):
    """Both radix plans of the x passes (fft_xplan 0: 8.2.8 / 8.4.8, 1:
    4.4.8 / 4.8.8), with the x transforms in launches of their own, against
    numpy: forward, inverse and the Poisson solve."""
    print(f"[fft xplan] n={n} prec={prec}", flush=True)
    rng = np.random.default_rng(950 + n + prec)
    x = rng.standard_normal((n, n, n))
    tol = 1e-12 if prec == 0 else 2e-6
    X = np.fft.rfftn(x)
    k = np.fft.fftfreq(n, 1.0 / n)
    kz, ky, kx = np.meshgrid(k, k, k[: n // 2 + 1], indexing="ij")
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    k2[0, 0, 0] = 1.0
    ph = -X / k2
    ph[0, 0, 0] = 0.0
    ref = np.fft.irfftn(ph, s=(n, n, n), axes=(0, 1, 2))
    with _solver(n, prec) as s:
        for plan in (1, 0):
            s.set_tuning(fft_xy=0, fft_xplan=plan)
            assert s.tuning()["fft_xplan"] == plan
            assert relerr(s.fft_r2c(x), X) < tol, plan
            assert relerr(s.fft_c2r(X), x * n ** 3) < tol, plan
            assert relerr(s.poisson_solve(x), ref) < (1e-12 if prec == 0 else 1e-5), plan
