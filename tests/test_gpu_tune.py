"""The autotuner leaves the state as it found it, and no launch shape
changes a result bit (the transform forms, fft_wave and fft_xy, change
rounding only; tests/test_gpu_fft.py checks every form)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _by_tag(t):
    o = np.argsort(t["tag"])
    return {k: v[o] for k, v in t.items()}


def test_autotune_restores_state_and_results_are_shape_independent():
    from mhit36_mixedp_amd import Solver
    print("[autotune] 48^3 multiphase + tracers", flush=True)
    kw = dict(nx=48, phiflag=1, partflag=1, sigma=0.5, dt=5e-4, poisson_precision=1)
    with Solver(**kw) as a, Solver(**kw) as b:
        for s in (a, b):
            s.init_random_velocity(9, kmax=4.0, urms=1.0)
            s.init_drops(3, 0.6, 9)
            s.init_tracers_uniform(4000, 9)
            s.step(2)
        t = a.autotune(steps=2)
        print(t, flush=True)
        # the transform form changes rounding (and is a no-op at 48^3; so
        # are the FFT widths, compiled with alternatives from 64 up); the
        # launch shapes change no bit
        b.set_tuning(fft_wave=t["fft_wave"], fft_ycols=t["fft_ycols"], fft_zcols=t["fft_zcols"],
                     fft_xrows=t["fft_xrows"], fft_xy=t["fft_xy"], fft_xplan=t["fft_xplan"])
        assert t["ms_per_step"] > 0 and t["fft_xy"] in (0, 1)
        assert t["ns_mz"] in (4, 8, 16, 32) and t["ac_mz"] in (4, 8, 16, 32) and t["xcd"] in (0, 1)
        assert t["ns_my"] in (8, 16, 32, 10) and t["ac_my"] in (8, 16, 32, 10) and t["fft_wave"] in (0, 1)
        assert a.tuning()["ns_mz"] == t["ns_mz"]
        # every field, the last projection's pressure (4) included
        for f in range(1, 10):
            assert np.array_equal(a.download_field(f), b.download_field(f)), f
        a.step(3)
        b.step(3)
        for f in (1, 2, 3, 4, 5):
            assert np.array_equal(a.download_field(f), b.download_field(f)), f
        ta, tb = _by_tag(a.download_tracers()), _by_tag(b.download_tracers())
        for k in ta:
            assert np.array_equal(ta[k], tb[k]), k
        assert a.diag() == b.diag()


@pytest.mark.parametrize("n,shape", [
    (64, dict(ns_mz=8, ac_mz=16, xcd=1, ns_my=16, ac_my=16)),
    (64, dict(ns_mz=64, ac_mz=128, xcd=0, ns_my=8, ac_my=16)),
    (64, dict(ns_mz=16, ac_mz=8, xcd=0, ns_my=16, ac_my=8)),
    # 4-plane chunks: warm-up planes longer than the chunk
    (64, dict(ns_mz=4, ac_mz=4, xcd=0, ns_my=8, ac_my=8)),
    (48, dict(ns_mz=4, ac_mz=4, xcd=1, ns_my=16, ac_my=16)),
    # 24 rows: the 16-row tiles end in a partial tile (the default is 8)
    (24, dict(ns_mz=8, ac_mz=8, xcd=0, ns_my=16, ac_my=16)),
    # x passes: no wavefront form, so every row-pair width is bit-neutral
    # (with the x and y transforms in separate launches, on both solvers)
    (64, dict(fft_xrows=4, fft_xy=0)),
    (96, dict(fft_xrows=16, fft_xy=0)),
    # 32 x 32 tiles, full and partial (48 = 32 + 16)
    (64, dict(ns_my=32, ac_my=32, ns_mz=16, ac_mz=16)),
    (96, dict(ns_my=16, ac_my=8)),
    (48, dict(ns_my=32, ac_my=32, ns_mz=8, ac_mz=4, xcd=1)),
    # 96 x 10 tiles (the default at 96^3): a partial last row of tiles
    # (96 = 9 x 10 + 6), and partial in x and y at 64 and 48
    (96, dict(ns_my=32, ac_my=32)),
    (96, dict(ns_my=10, ac_my=16, ns_mz=12, ac_mz=24, xcd=0)),
    (64, dict(ns_my=10, ac_my=10, ns_mz=16, ac_mz=8)),
    (48, dict(ns_my=10, ac_my=10, ns_mz=8, ac_mz=4, xcd=1)),
])
def test_set_tuning_is_bitwise_neutral(n, shape):
    from mhit36_mixedp_amd import Solver
    print(f"[tuning] n={n} {shape}", flush=True)
    kw = dict(nx=n, phiflag=1, sigma=0.5, dt=5e-4, poisson_precision=1)
    with Solver(**kw) as a, Solver(**kw) as b:
        a.set_tuning(**shape)
        # the transform forms change rounding: the same on both
        b.set_tuning(**{k: v for k, v in shape.items() if k in ("fft_wave", "fft_xy")})
        for s in (a, b):
            s.init_random_velocity(4, kmax=4.0, urms=1.0)
            s.init_drops(2, 0.7, 4)
            s.step(4)
        for f in (1, 2, 3, 4, 5):
            assert np.array_equal(a.download_field(f), b.download_field(f)), f


def test_mhit_tune_overrides_and_ignores_bad_values(monkeypatch):
    """MHIT_TUNE is read at context creation: valid keys override the
    size-based shapes, unknown keys, non-positive chunks and tile rows
    other than 8 or 16 are ignored."""
    from mhit36_mixedp_amd import Solver
    print("[tuning] MHIT_TUNE parsing", flush=True)
    with Solver(nx=32, phiflag=1, sigma=0.5) as s:
        base = s.tuning()
    monkeypatch.setenv("MHIT_TUNE", "ns_mz=16,ac_mz=8,ns_my=16,ac_my=16,xcd=1,sort=3,fft=0,xy=0")
    with Solver(nx=32, phiflag=1, sigma=0.5) as s:
        t = s.tuning()
    assert (t["ns_mz"], t["ac_mz"], t["ns_my"], t["ac_my"], t["xcd"], t["sort_every"], t["fft_wave"],
            t["fft_xy"]) == (16, 8, 16, 16, 1, 3, 0, 0)
    monkeypatch.setenv("MHIT_TUNE", "ns_mz=0,ac_mz=-4,ns_my=12,ac_my=4,bogus=7,sort=-1")
    with Solver(nx=32, phiflag=1, sigma=0.5) as s:
        t = s.tuning()
    for k in ("ns_mz", "ac_mz", "ns_my", "ac_my", "xcd", "sort_every"):
        assert t[k] == base[k], (k, t[k], base[k])


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
@pytest.mark.parametrize("n", [64, 96, 128])
def test_fft_widths_are_bitwise_neutral_in_one_form(n, prec):
    """Every compiled FFT launch width gives the same bits as the default
    when both use the register-direct form (the wavefront form exists only
    at eight columns)."""
    from mhit36_mixedp_amd import Solver
    print(f"[tuning] FFT widths n={n} prec={prec}", flush=True)
    kw = dict(nx=n, phiflag=1, sigma=0.5, dt=5e-4, poisson_precision=prec)
    with Solver(**kw) as a, Solver(**kw) as b:
        base = b.tuning()
        alt = dict(fft_ycols=base["fft_ycols"] // 2,
                   fft_zcols=base["fft_zcols"] * 2 if prec == 1 else base["fft_zcols"] // 2,
                   fft_xrows=base["fft_xrows"] // 2)
        a.set_tuning(fft_wave=0, fft_xy=0, **alt)
        b.set_tuning(fft_wave=0, fft_xy=0)
        t = a.tuning()
        assert all(t[k] == v for k, v in alt.items()), (t, alt)
        for s in (a, b):
            s.init_random_velocity(5, kmax=4.0, urms=1.0)
            s.init_drops(2, 0.7, 5)
            s.step(4)
        for f in (1, 2, 3, 4, 5):
            assert np.array_equal(a.download_field(f), b.download_field(f)), f


def test_fft_width_validation():
    """Widths that are not compiled for the size are rejected by
    mhit_set_tuning; 0 restores the default."""
    from mhit36_mixedp_amd import MhitError, Solver
    with Solver(nx=64, poisson_precision=1) as s:
        base = s.tuning()
        assert (base["fft_ycols"], base["fft_zcols"], base["fft_xrows"]) == (16, 8, 8)
        for bad in (dict(fft_ycols=32), dict(fft_zcols=4), dict(fft_xrows=3)):
            with pytest.raises(MhitError):
                s.set_tuning(**bad)
        s.set_tuning(fft_ycols=8)
        assert s.tuning()["fft_ycols"] == 8
        s.set_tuning(fft_ycols=0)
        assert s.tuning()["fft_ycols"] == 16
    with Solver(nx=32, poisson_precision=1) as s:   # no alternatives below 64
        with pytest.raises(MhitError):
            s.set_tuning(fft_xrows=4)


@pytest.mark.parametrize("n,prec,compiled", [(24, 1, False), (32, 1, True), (64, 0, True), (96, 0, True),
                                             (128, 1, True), (128, 0, False), (256, 1, False)])
def test_fft_xy_default_and_validation(n, prec, compiled):
    """The one-launch x/y transforms are the default exactly where they are
    compiled (n = 32 ... 128 FP32, 32 ... 96 FP64); elsewhere the
    context reports 0 whatever was asked; values other than 0 and 1 are
    rejected."""
    from mhit36_mixedp_amd import MhitError, Solver
    print(f"[tuning] fft_xy n={n} prec={prec}", flush=True)
    with Solver(nx=n, poisson_precision=prec) as s:
        assert s.tuning()["fft_xy"] == int(compiled)
        s.set_tuning(fft_xy=0)
        assert s.tuning()["fft_xy"] == 0
        s.set_tuning(fft_xy=1)
        assert s.tuning()["fft_xy"] == int(compiled)
        with pytest.raises(MhitError):
            s.set_tuning(fft_xy=2)


@pytest.mark.parametrize("n,two", [(64, False), (96, False), (128, True), (256, True), (192, False)])
def test_fft_xplan_validation(n, two):
    """The x passes have a second radix plan only where it differs from the
    default (n = 128, 256, 384, 512); elsewhere the context reports 0."""
    from mhit36_mixedp_amd import MhitError, Solver
    print(f"[tuning] fft_xplan n={n}", flush=True)
    with Solver(nx=n, poisson_precision=1) as s:
        assert s.tuning()["fft_xplan"] == 0
        s.set_tuning(fft_xplan=1)
        assert s.tuning()["fft_xplan"] == int(two)
        with pytest.raises(MhitError):
            s.set_tuning(fft_xplan=-1)
