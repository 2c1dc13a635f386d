"""An ensemble of boxes stepped together equals each box stepped alone.

Boxes of different sizes and physics (single-phase, multiphase with
tracers) run concurrently on their own streams; each must reproduce, bit
for bit, the same box advanced by a standalone Solver."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BOXES = [
    dict(nx=32, phiflag=1, partflag=1, sigma=0.5),
    dict(nx=24, phiflag=0, partflag=0),
    dict(nx=48, phiflag=1, partflag=0, sigma=0.5),
    dict(nx=16, phiflag=0, partflag=1),
]


def _params(kw):
    from mhit36_mixedp_amd import default_params
    return default_params(poisson_precision=1, dt=5e-4, mu=0.006, **kw)


def _init(s, seed, kw):
    s.init_random_velocity(seed, kmax=4.0, urms=1.0)
    if kw["phiflag"]:
        s.init_drops(2, 0.6, seed)
    if kw["partflag"]:
        s.init_tracers_uniform(1500, seed)


def test_ensemble_equals_solo():
    from mhit36_mixedp_amd import Ensemble, Solver
    print("[ensemble] 4 boxes 32/24/48/16", flush=True)
    params = [_params(kw) for kw in BOXES]
    with Ensemble(params) as ens:
        solos = [Solver(_params(kw)) for kw in BOXES]
        for i, kw in enumerate(BOXES):
            _init(ens.boxes[i], 11 + i, kw)
            _init(solos[i], 11 + i, kw)
        for chunk in (1, 3, 4):
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


def test_ensemble_cfl_stop_isolated():
    """A box stopped on cou > 7 stays frozen; the other boxes go on."""
    from mhit36_mixedp_amd import CFLStop, Ensemble
    print("[ensemble] CFL isolation", flush=True)
    good = _params(dict(nx=16, phiflag=0, partflag=0))
    bad = _params(dict(nx=16, phiflag=0, partflag=0))
    bad.dt = 0.5
    with Ensemble([good, bad]) as ens:
        ens.boxes[0].init_taylor_green()
        ens.boxes[1].init_random_velocity(5, kmax=3.0, urms=20.0)
        with pytest.raises(CFLStop):
            ens.step(3)
        assert ens.boxes[1].diag()["step"] == 1
        assert ens.boxes[0].diag()["step"] == 3
        with pytest.raises(CFLStop):
            ens.step(2)
        assert ens.boxes[1].diag()["step"] == 1
        assert ens.boxes[0].diag()["step"] == 5
