"""The tracer re-ordering: a stable sort of the tracers by the index of
their cell, (k n + j) n + i (a hand-written LSD radix sort, 8 bits per
pass).

The step after an upload re-orders first.  From rest the first (Euler)
step moves no tracer, so after one step the tracer columns must be exactly
the uploaded ones permuted by numpy's stable argsort of the cell codes:
equal codes keep their upload order.  The sizes cover a partial last tile
(tiles are 4096 tracers), a single tile, two, three and four digit passes
(the passes cover 3 ceil(log2 n) bits), and many tracers per cell."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cell_codes(t, n):
    dx = 2.0 * np.pi / n
    c = [np.mod(np.floor(t[a] / dx).astype(np.int64), n) for a in ("x", "y", "z")]
    return (c[2] * n + c[1]) * n + c[0]


@pytest.mark.parametrize("n,ntr", [(8, 3000), (16, 4096), (64, 1_000_003), (256, 262_145), (512, 4_194_304)])
def test_reorder_is_stable_cell_order(n, ntr):
    from mhit36_mixedp_amd import Solver
    print(f"[tracer sort] n={n} tracers={ntr}", flush=True)
    rng = np.random.default_rng(n + ntr)
    x, y, z = (rng.random(ntr) * (2.0 * np.pi) for _ in range(3))
    # a tenth of the tracers on a handful of cells: long runs of equal codes
    k = ntr // 10
    x[:k] = (rng.integers(0, 3, k) + 0.5) * (2.0 * np.pi / n)
    y[:k] = (rng.integers(0, 2, k) + 0.25) * (2.0 * np.pi / n)
    z[:k] = (rng.integers(0, 2, k) + 0.75) * (2.0 * np.pi / n)
    with Solver(nx=n, partflag=1, phiflag=0, poisson_precision=0) as s:
        zero = np.zeros((n, n, n))
        s.set_velocity(zero, zero, zero)
        s.upload_tracers(x, y, z)
        t0 = s.download_tracers()
        s.step(1)
        t1 = s.download_tracers()
    order = np.argsort(_cell_codes(t0, n), kind="stable")
    for col in ("tag", "x", "y", "z"):
        assert np.array_equal(t1[col], t0[col][order]), col
    assert np.all(np.diff(_cell_codes(t1, n)) >= 0)
