"""An ensemble of boxes stepped together equals each box stepped alone.

Boxes of different sizes, physics (single-phase, multiphase, tracers) and
Poisson precisions run together, batched (boxes of one size share every
launch of a pass, grouped by phase flag and precision) or one graph branch
per box; each must reproduce, bit for bit, the same box advanced by a
standalone Solver."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BOXES = [
    dict(nx=32, phiflag=1, partflag=1, sigma=0.5),
    dict(nx=24, phiflag=0, partflag=0),
    dict(nx=48, phiflag=1, partflag=0, sigma=0.5),
    dict(nx=16, phiflag=0, partflag=1),
    # more boxes of the same sizes: batched launches over several boxes,
    # mixed phase flags and precisions within a size
    dict(nx=32, phiflag=0, partflag=0),
    dict(nx=32, phiflag=1, partflag=0, sigma=0.5),
    dict(nx=32, phiflag=1, partflag=0, sigma=0.5, prec=0),
    dict(nx=24, phiflag=0, partflag=1),
    dict(nx=48, phiflag=1, partflag=1, sigma=0.5),
]


def _params(kw):
    from mhit36_mixedp_amd import default_params
    kw = dict(kw)
    prec = kw.pop("prec", 1)
    return default_params(poisson_precision=prec, dt=5e-4, mu=0.006, **kw)


def _init(s, seed, kw):
    s.init_random_velocity(seed, kmax=4.0, urms=1.0)
    if kw["phiflag"]:
        s.init_drops(2, 0.6, seed)
    if kw["partflag"]:
        s.init_tracers_uniform(1500, seed)


@pytest.mark.parametrize("batched", [True, False])
def test_ensemble_equals_solo(batched):
    from mhit36_mixedp_amd import Ensemble, Solver
    print(f"[ensemble] {len(BOXES)} boxes of 16..48, batched={batched}", flush=True)
    params = [_params(kw) for kw in BOXES]
    with Ensemble(params, batched=batched) as ens:
        solos = [Solver(_params(kw)) for kw in BOXES]
        for i, kw in enumerate(BOXES):
            _init(ens.boxes[i], 11 + i, kw)
            _init(solos[i], 11 + i, kw)
        for chunk in (1, 3, 4):
            if chunk == 3:
                # a launch shape of its own for one box of size 32: its
                # NS pass leaves the shared launch (results unchanged)
                ens.boxes[4].set_tuning(ns_mz=16)
            if chunk == 4:
                # the x and y transforms in two launches for another box of
                # size 32: its Poisson passes leave the shared launches
                # (rounding changes, so on its solo too)
                ens.boxes[5].set_tuning(fft_xy=0)
                solos[5].set_tuning(fft_xy=0)
            ens.step(chunk)
            for s in solos:
                s.step(chunk)
            for i, kw in enumerate(BOXES):
                a, b = ens.boxes[i], solos[i]
                fields = (1, 2, 3, 4, 5) if kw["phiflag"] else (1, 2, 3, 4)
                for f in fields:
                    assert np.array_equal(a.download_field(f), b.download_field(f)), (i, f)
                if kw["partflag"]:
                    ta, tb = a.download_tracers(), b.download_tracers()
                    for k in ("tag", "x", "y", "z", "up"):
                        assert np.array_equal(ta[k], tb[k]), (i, k)
                assert a.diag() == b.diag(), i
        for s in solos:
            s.close()


@pytest.mark.parametrize("batched", [True, False])
def test_ensemble_cfl_stop_isolated(batched):
    """A box stopped on cou > 7 stays frozen; the other boxes go on (in the
    batched form both boxes share every launch).  The frozen box goes
    through three no-op steps inside the first call (an odd number of swaps
    of the AB2 ping-pong buffers) and keeps the state, AB2 history
    included, of the same box stopped alone."""
    from mhit36_mixedp_amd import CFLStop, Ensemble, Solver
    print(f"[ensemble] CFL isolation, batched={batched}", flush=True)
    good = _params(dict(nx=16, phiflag=0, partflag=0))
    bad = _params(dict(nx=16, phiflag=1, partflag=0, sigma=0.5))
    bad.dt = 0.5
    with Ensemble([good, bad], batched=batched) as ens, Solver(bad) as solo:
        ens.boxes[0].init_taylor_green()
        for s in (ens.boxes[1], solo):
            s.init_random_velocity(5, kmax=3.0, urms=20.0)
            s.init_drops(2, 0.6, 5)
        with pytest.raises(CFLStop):
            ens.step(4)
        with pytest.raises(CFLStop):
            solo.step(1)
        assert ens.boxes[1].diag()["step"] == 1
        assert ens.boxes[0].diag()["step"] == 4
        with pytest.raises(CFLStop):
            ens.step(2)
        assert ens.boxes[1].diag()["step"] == 1
        assert ens.boxes[0].diag()["step"] == 6
        for f in range(1, 10):
            assert np.array_equal(ens.boxes[1].download_field(f), solo.download_field(f),
                                  equal_nan=True), f


def _ns_workgroups(boxes):
    total = 0
    for b in boxes:
        n, t = b.n, b.tuning()
        total += -(-n // 64) * -(-n // t["ns_my"]) * -(-n // t["ns_mz"])
    return total


def test_batched_shapes_fill_the_gpu():
    """Batched boxes get z chunks halved from min(n, 128) (never below 8
    planes) until the NS launches of one step hold about four workgroups
    per CU (on a 256-CU MI355X: 32 planes for 16 boxes of 128^3, 16 for 8
    boxes); unbatched boxes get the shapes of a box alone.  Checked without
    the CU count: half the boxes get half the chunk and the same number of
    workgroups."""
    from mhit36_mixedp_amd import Ensemble, Solver
    print("[ensemble] batched launch shapes, 16 and 8 boxes of 128^3", flush=True)
    chosen = {}
    for nbox in (16, 8):
        params = [_params(dict(nx=128, phiflag=i % 2, partflag=0, sigma=0.5)) for i in range(nbox)]
        with Ensemble(params) as ens:
            t = [b.tuning() for b in ens.boxes]
            assert len({(x["ns_mz"], x["ac_mz"], x["ns_my"]) for x in t}) == 1
            zc = t[0]["ns_mz"]
            assert zc == t[0]["ac_mz"] and t[0]["ns_my"] == 16 and 8 <= zc <= 128 and 128 % zc == 0
            chosen[nbox] = (zc, _ns_workgroups(ens.boxes))
            if nbox == 8:
                with Solver(_params(dict(nx=128, phiflag=0, partflag=0))) as solo:
                    alone = solo.tuning()
                ens.set_batched(False)
                for b in ens.boxes:
                    x = b.tuning()
                    assert (x["ns_mz"], x["ns_my"], x["ac_my"]) == (alone["ns_mz"], alone["ns_my"], alone["ac_my"])
                ens.set_batched(True)
                assert all(b.tuning()["ns_mz"] == zc for b in ens.boxes)
    (z16, w16), (z8, w8) = chosen[16], chosen[8]
    print(chosen, flush=True)
    if z16 < 128 and z8 > 8:
        # halving stops at the first workgroup count that reaches the
        # target: the same count for 16 boxes and for 8
        assert z8 * 2 == z16 and w8 == w16


def test_ensemble_autotune_restores_state_and_matches_solo():
    """mhit_ensemble_autotune leaves every box's state as it found it, and
    the tuned ensemble still equals each box stepped alone with the same
    transform form and FFT widths (the launch shapes change no bit)."""
    from mhit36_mixedp_amd import Ensemble, Solver
    kws = [dict(nx=64, phiflag=1, partflag=1, sigma=0.5), dict(nx=64, phiflag=0, partflag=0),
           dict(nx=96, phiflag=1, partflag=0, sigma=0.5), dict(nx=96, phiflag=0, partflag=1)]
    print("[ensemble] autotune, 2 x 64^3 + 2 x 96^3", flush=True)

    def fields(kw):
        return (1, 2, 3, 4, 5, 6, 7, 8, 9) if kw["phiflag"] else (1, 2, 3, 4, 6, 7, 8)

    def state(s, kw):
        st = {f: s.download_field(f) for f in fields(kw)}
        if kw["partflag"]:
            st.update({"t_" + k: v for k, v in s.download_tracers().items()})
        return st

    with Ensemble([_params(kw) for kw in kws]) as ens:
        solos = [Solver(_params(kw)) for kw in kws]
        for i, kw in enumerate(kws):
            _init(ens.boxes[i], 40 + i, kw)
            _init(solos[i], 40 + i, kw)
        ens.step(2)
        before = [state(b, kw) for b, kw in zip(ens.boxes, kws)]
        ms = ens.autotune(steps=2)
        assert ms > 0
        for i, kw in enumerate(kws):
            after = state(ens.boxes[i], kw)
            for k in before[i]:
                assert np.array_equal(before[i][k], after[k]), (i, k)
        for i, s in enumerate(solos):
            t = ens.boxes[i].tuning()
            print(i, t, flush=True)
            s.step(2)
            s.set_tuning(fft_wave=t["fft_wave"], fft_ycols=t["fft_ycols"], fft_zcols=t["fft_zcols"],
                         fft_xrows=t["fft_xrows"], fft_xy=t["fft_xy"], fft_xplan=t["fft_xplan"])
        ens.step(3)
        for s in solos:
            s.step(3)
        for i, kw in enumerate(kws):
            a, b = state(ens.boxes[i], kw), state(solos[i], kw)
            for k in a:
                assert np.array_equal(a[k], b[k]), (i, k)
            assert ens.boxes[i].diag() == solos[i].diag(), i
        for s in solos:
            s.close()
