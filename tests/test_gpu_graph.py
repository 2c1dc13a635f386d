"""Captured-graph timesteps equal eagerly launched ones, bit for bit.

The graph path (default) is replayed many times with eager work between
the replays (downloads, diagnostics, re-uploads) and across the buffer
changes that select a different captured graph: the phi ping-pong of the
fused ACDI pass and the tracer re-ordering every few steps."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(monkeypatch, n=32, **kw):
    from mhit36_mixedp_amd import Solver
    kw = dict(nx=n, poisson_precision=1, phiflag=1, partflag=1, dt=5e-4, mu=0.006, sigma=0.5, **kw)
    monkeypatch.setenv("MHIT_TUNE", "sort=4")
    monkeypatch.setenv("MHIT_GRAPH", "1")
    g = Solver(**kw)
    monkeypatch.setenv("MHIT_GRAPH", "0")
    e = Solver(**kw)
    for s in (g, e):
        s.init_random_velocity(7, kmax=4.0, urms=1.0)
        s.init_drops(3, 0.6, 7)
        s.init_tracers_uniform(3000, 7)
    return g, e


def _same(g, e):
    for f in (1, 2, 3, 4, 5):
        assert np.array_equal(g.download_field(f), e.download_field(f)), f
    tg, te = g.download_tracers(), e.download_tracers()
    for k in ("tag", "x", "y", "z", "up"):
        assert np.array_equal(tg[k], te[k]), k


def test_graph_replays_equal_eager(monkeypatch):
    print("[graph vs eager] 32^3 multiphase + tracers", flush=True)
    g, e = _pair(monkeypatch)
    for chunk in (1, 2, 3, 5, 1, 4):
        g.step(chunk)
        e.step(chunk)
        _same(g, e)
        assert g.diag() == e.diag()
    # eager work that changes the state between replays
    u = g.download_field(1)
    for s in (g, e):
        s.upload_field(1, u * 0.5)
        s.init_tracers_uniform(2000, 9)      # new tracer count: graphs re-captured
    g.step(6)
    e.step(6)
    _same(g, e)
    g.close()
    e.close()


def test_graph_cfl_stop(monkeypatch):
    """The CFL stop inside a captured step: later replays do nothing."""
    from mhit36_mixedp_amd import CFLStop, Solver
    monkeypatch.setenv("MHIT_GRAPH", "1")
    with Solver(nx=16, dt=0.5) as s:
        s.init_random_velocity(5, kmax=3.0, urms=20.0)
        with pytest.raises(CFLStop):
            s.step(4)
        assert s.diag()["step"] == 1


@pytest.mark.parametrize("graph", ["1", "0"], ids=["graph", "eager"])
def test_cfl_stop_mid_call_keeps_state(monkeypatch, graph):
    """Steps queued behind the CFL stop change nothing: the state (phi's
    ping-pong buffers and the AB2 history, fields 6-9, included) is the one
    after the stopping step.  step(4) queues three no-op steps behind the
    stop, an odd number of swaps of the AB2 ping-pong buffers."""
    from mhit36_mixedp_amd import CFLStop, Solver
    monkeypatch.setenv("MHIT_GRAPH", graph)
    kw = dict(nx=16, dt=0.5, phiflag=1, sigma=0.5)
    with Solver(**kw) as a, Solver(**kw) as b:
        for s in (a, b):
            s.init_random_velocity(5, kmax=3.0, urms=20.0)
            s.init_drops(2, 0.6, 5)
        with pytest.raises(CFLStop):
            a.step(4)
        with pytest.raises(CFLStop):
            b.step(1)
        for f in range(1, 10):
            assert np.array_equal(a.download_field(f), b.download_field(f), equal_nan=True), f
        assert a.diag()["step"] == 1
        # a resumed run continues from the history of the last step taken
        # (it may stop again: both runs then stop alike)
        for s in (a, b):
            s.set_ab2_state(1.5, 0.5, 1.0)
            s.upload_field(1, np.zeros((16, 16, 16)))
            s.upload_field(2, np.zeros((16, 16, 16)))
            s.upload_field(3, np.zeros((16, 16, 16)))
            try:
                s.step(1)
            except CFLStop:
                pass
        assert a.diag()["step"] == b.diag()["step"] == 2
        for f in range(1, 10):
            assert np.array_equal(a.download_field(f), b.download_field(f), equal_nan=True), f
