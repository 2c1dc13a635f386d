"""GPU drop statistics (union-find labelling with atomic links) against the
oracle's flood fill: identical components (count, cells, first cell) and
centres, velocities, volumes and d_eq to round-off of the summation order."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _compare(g, o):
    assert g["count"] == o["count"]
    assert np.array_equal(g["cells"], o["cells"])
    assert np.array_equal(g["first_cell"], o["first_cell"])
    for k in ("volume", "d_eq"):
        assert np.allclose(g[k], o[k], rtol=1e-13, atol=0)
    for k in ("u", "v", "w"):
        assert np.allclose(g[k], o[k], rtol=1e-10, atol=1e-13)
    for k in ("x", "y", "z"):
        d = np.abs(g[k] - o[k])
        assert np.all(np.minimum(d, 2 * np.pi - d) < 1e-9), k


def _solver(n):
    from mhit36_mixedp_amd import Solver
    return Solver(nx=n, phiflag=1, poisson_precision=1)


@pytest.mark.parametrize("conn", [6, 26])
@pytest.mark.parametrize("n", [24, 48, 96])
def test_percolating_random_field(oracle_mod, n, conn):
    """Random field above and near the site-percolation threshold: a giant
    periodic cluster plus many small ones stress concurrent unions."""
    print(f"[drops] random n={n} conn={conn}", flush=True)
    rng = np.random.default_rng(n + conn)
    phi = rng.uniform(0, 1, (n, n, n))
    u, v, w = (rng.standard_normal((n, n, n)) for _ in range(3))
    with _solver(n) as s:
        s.upload_field(5, phi)
        s.set_velocity(u, v, w)
        for thr in (0.5, 0.75):
            g = s.drop_stats(threshold=thr, connectivity=conn)
            o = oracle_mod.drop_stats(phi, u, v, w, threshold=thr, connectivity=conn)
            _compare(g, o)


def test_seeded_drops_after_steps(oracle_mod):
    """Drops placed by init_drops, advected a few steps: the GPU statistics
    of the live state equal the oracle's on the downloaded fields."""
    print("[drops] 64^3 seeded drops", flush=True)
    from mhit36_mixedp_amd import Solver
    with Solver(nx=64, phiflag=1, poisson_precision=1, dt=5e-4, sigma=0.5) as s:
        s.init_random_velocity(3, kmax=4.0, urms=1.0)
        s.init_drops(12, 0.45, 3)
        s.step(5)
        phi = s.download_field(5)
        u, v, w = s.velocity()
        g = s.drop_stats()
        o = oracle_mod.drop_stats(phi, u, v, w)
        _compare(g, o)
        assert g["count"] == 12
        assert np.allclose(g["d_eq"], 0.9, rtol=0.08)
        # the labelling leaves the state untouched (it uses scratch memory)
        assert np.array_equal(s.download_field(5), phi)
        s.step(1)


def test_max_drops_truncates():
    print("[drops] truncation", flush=True)
    n = 32
    phi = np.zeros((n, n, n))
    phi[::4, ::4, ::4] = 1.0                      # 512 isolated cells
    with _solver(n) as s:
        s.upload_field(5, phi)
        g = s.drop_stats(max_drops=10)
        assert g["count"] == 512 and len(g["cells"]) == 10
        assert list(g["first_cell"]) == [0, 4, 8, 12, 16, 20, 24, 28, 128, 132]
