"""Pins of the oracle's drop statistics (post/unionfind_periodic/main.f90,
post/DropCountSize/calculate_deq.f90): closed forms for balls and cubes,
periodic wrap, 6- versus 26-point connectivity, and an independent
brute-force labelling on small random masks."""
import itertools

import numpy as np
import pytest

L = 2 * np.pi


def _stats(oracle_mod, phi, **kw):
    return oracle_mod.drop_stats(phi, **kw)


def _centres(n):
    return (np.arange(n) + 0.5) * (L / n)


def _ball(n, c, r):
    x = _centres(n)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    d2 = sum(np.minimum(np.abs(a - b), L - np.abs(a - b)) ** 2 for a, b in zip((X, Y, Z), c))
    return (d2 <= r * r).astype(float)


def test_single_ball_volume_centre_deq(oracle_mod):
    n, r = 48, 1.0
    c = (2.0, 3.5, 4.2)
    phi = _ball(n, c, r)
    s = _stats(oracle_mod, phi)
    assert s["count"] == 1
    dx = L / n
    assert s["cells"][0] == phi.sum()
    assert s["volume"][0] == pytest.approx(phi.sum() * dx ** 3, rel=1e-15)
    assert s["volume"][0] == pytest.approx(4 / 3 * np.pi * r ** 3, rel=0.03)
    assert s["d_eq"][0] == pytest.approx(2 * r, rel=0.01)
    # circular mean of a symmetric ball's cell centres: near the ball centre
    for k, a in zip(("x", "y", "z"), c):
        assert abs(s[k][0] - a) < 0.5 * dx


def test_ball_across_periodic_corner(oracle_mod):
    n = 32
    phi = _ball(n, (0.0, 0.0, 0.0), 0.8)
    s = _stats(oracle_mod, phi)
    assert s["count"] == 1 and s["first_cell"][0] == 0
    for k in ("x", "y", "z"):
        d = min(s[k][0], L - s[k][0])
        assert d < L / n


def test_edge_and_corner_contact_connectivity(oracle_mod):
    n = 16
    phi = np.zeros((n, n, n))
    phi[2:5, 2:5, 2:5] = 1.0          # cube A, cells z,y,x in 2..4
    phi[5:8, 5:8, 2:5] = 1.0          # cube B touches A along an x-edge (dz=dy=1)
    phi[10:12, 10:12, 10:12] = 1.0    # cube C
    phi[12:14, 12:14, 12:14] = 1.0    # cube D touches C at a corner only
    s6 = _stats(oracle_mod, phi, connectivity=6)
    s26 = _stats(oracle_mod, phi, connectivity=26)
    assert s6["count"] == 4 and s26["count"] == 2
    assert list(s6["cells"]) == [27, 27, 8, 8]
    assert list(s26["cells"]) == [54, 16]
    # numbering follows the first cell in (z, y, x) scan order
    assert list(s6["first_cell"]) == sorted(s6["first_cell"])


def test_wrap_joins_faces(oracle_mod):
    n = 12
    phi = np.zeros((n, n, n))
    phi[3, 4, 0] = phi[3, 4, n - 1] = 1.0     # neighbours through the x face
    phi[0, 7, 7] = phi[n - 1, 7, 7] = 1.0     # through the z face
    phi[6, 0, 2] = phi[6, n - 1, 3] = 1.0     # y face + x edge: 26 only
    s6 = _stats(oracle_mod, phi, connectivity=6)
    s26 = _stats(oracle_mod, phi, connectivity=26)
    assert s6["count"] == 4 and s26["count"] == 3


def test_threshold_is_inclusive(oracle_mod):
    n = 8
    phi = np.zeros((n, n, n))
    phi[1, 1, 1] = 0.5
    phi[5, 5, 5] = 0.5 - 1e-12
    assert _stats(oracle_mod, phi, threshold=0.5)["count"] == 1


def test_mean_velocity(oracle_mod):
    n = 24
    phi = _ball(n, (3.0, 3.0, 3.0), 1.2)
    x = _centres(n)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    u = 0.3 + 0 * X
    v = np.sin(X)
    w = Z.copy()
    s = _stats(oracle_mod, phi, u=u, v=v, w=w)
    m = phi > 0
    assert s["u"][0] == pytest.approx(0.3, rel=1e-14)
    assert s["v"][0] == pytest.approx(v[m].mean(), rel=1e-12)
    assert s["w"][0] == pytest.approx(w[m].mean(), rel=1e-12)


def _brute_components(mask, conn):
    """Independent labelling: repeated min-label sweeps to a fixed point,
    as the reference does (post/unionfind_periodic/main.f90:193-285)."""
    n = mask.shape[0]
    lab = np.where(mask, np.arange(n ** 3).reshape(mask.shape), -1)
    offs = [o for o in itertools.product((-1, 0, 1), repeat=3)
            if any(o) and (conn == 26 or sum(map(abs, o)) == 1)]
    while True:
        new = lab.copy()
        for o in offs:
            sh = np.roll(lab, shift=o, axis=(0, 1, 2))
            ok = mask & (sh >= 0)
            new[ok] = np.minimum(new[ok], sh[ok])
        if np.array_equal(new, lab):
            break
        lab = new
    roots = np.unique(lab[mask])
    return [(int(r), int((lab == r).sum())) for r in roots]


@pytest.mark.parametrize("conn", [6, 26])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_masks_against_brute_force(oracle_mod, conn, seed):
    n = 7
    phi = np.random.default_rng(seed).uniform(0, 1, (n, n, n)) ** 2
    s = _stats(oracle_mod, phi, connectivity=conn, threshold=0.5)
    ref = _brute_components(phi >= 0.5, conn)
    assert s["count"] == len(ref)
    assert [(int(a), int(b)) for a, b in zip(s["first_cell"], s["cells"])] == ref
