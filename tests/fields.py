"""numpy generators of test fields (C order [z][y][x])."""
import numpy as np

L = 2 * np.pi


def divfree(n, seed, kmax=None):
    """Random velocity whose staggered forward-difference divergence
    sum_d (u_d(+e_d) - u_d) vanishes to roundoff (projection with the
    symbols c_d = exp(2 pi i m_d/n) - 1)."""
    rng = np.random.default_rng(seed)
    U = [np.fft.fftn(rng.standard_normal((n, n, n))) for _ in range(3)]
    k = np.fft.fftfreq(n, 1.0 / n)
    KZ, KY, KX = np.meshgrid(k, k, k, indexing="ij")
    if kmax is not None:
        mask = (KX ** 2 + KY ** 2 + KZ ** 2) <= kmax ** 2
        U = [a * mask for a in U]
    c = [np.exp(2j * np.pi * K / n) - 1 for K in (KX, KY, KZ)]
    dot = sum(U[d] * c[d] for d in range(3))
    cc = sum(np.abs(c[d]) ** 2 for d in range(3))
    cc[0, 0, 0] = 1.0
    U = [U[d] - dot * np.conj(c[d]) / cc for d in range(3)]
    out = [np.real(np.fft.ifftn(a)) for a in U]
    s = np.sqrt(sum((a ** 2).mean() for a in out) / 3)
    return [np.ascontiguousarray(a / s) for a in out]


def fdiv(u, v, w):
    """Forward-difference divergence (index units), main.f90:841-843."""
    return (np.roll(u, -1, 2) - u) + (np.roll(v, -1, 1) - v) + (np.roll(w, -1, 0) - w)


def lap7(p, dx):
    """7-point Laplacian with periodic wrap."""
    out = -6.0 * p
    for ax in range(3):
        out = out + np.roll(p, 1, ax) + np.roll(p, -1, ax)
    return out / dx ** 2
