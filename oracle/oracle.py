"""ctypes wrapper of the CPU oracle (oracle/libmhit_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from mhit36_mixedp_amd._abi import Diag, Params, default_params  # noqa: E402

LIB_PATH = os.path.join(HERE, "libmhit_oracle.so")

NS_CONV, NS_VISC, NS_FORCE = 1, 2, 4
PHI_CONV, PHI_DIFF, PHI_SHARPEN = 1, 2, 4

_P = ctypes.POINTER
_dp = _P(ctypes.c_double)
_fp = _P(ctypes.c_float)
_vp = ctypes.c_void_p

_SIGS = [
    ("oracle_create", _vp, [_P(Params)]),
    ("oracle_destroy", None, [_vp]),
    ("oracle_set_field", ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    ("oracle_get_field", ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    ("oracle_set_tracers", ctypes.c_int, [_vp, ctypes.c_int64, _dp, _dp, _dp]),
    ("oracle_get_tracers", ctypes.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp]),
    ("oracle_set_ab2_state", ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    ("oracle_step", ctypes.c_int, [_vp, ctypes.c_int]),
    ("oracle_diag", ctypes.c_int, [_vp, _P(Diag)]),
    ("oracle_set_threads", ctypes.c_int, [ctypes.c_int]),
    ("oracle_ns_rhs", ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp]),
    ("oracle_acdi_rhs", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, _dp]),
    ("oracle_csf", ctypes.c_int, [_vp, _dp, _dp, _dp]),
    ("oracle_divergence", ctypes.c_int, [_vp, _dp]),
    ("oracle_interp", ctypes.c_int, [_vp, _dp, _dp, _dp]),
    ("oracle_poisson", ctypes.c_int, [_vp, _dp, _dp]),
    ("oracle_fft1d", None, [ctypes.c_int, _dp, _dp, ctypes.c_int]),
    ("oracle_fft3d", None, [ctypes.c_int, _dp, _dp, ctypes.c_int]),
    ("oracle_drop_stats", ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_double,
                                          ctypes.c_int, ctypes.c_int64, _dp, _P(ctypes.c_int64)]),
    ("oracle_fft1d_f32", None, [ctypes.c_int, _fp, _fp, ctypes.c_int]),
]

_lib = None


def build() -> None:
    """Compile the oracle with its Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def set_threads(n: int) -> int:
    return lib().oracle_set_threads(int(n))


def fft1d(x: np.ndarray, sign: int = -1) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    lib().oracle_fft1d(x.size, x.ctypes.data_as(_dp), out.ctypes.data_as(_dp), sign)
    return out


def fft1d_f32(x: np.ndarray, sign: int = -1) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.complex64)
    out = np.empty_like(x)
    lib().oracle_fft1d_f32(x.size, x.ctypes.data_as(_fp), out.ctypes.data_as(_fp), sign)
    return out


def drop_stats(phi: np.ndarray, u=None, v=None, w=None, threshold: float = 0.5,
               connectivity: int = 6, max_drops: int = 1 << 16) -> dict:
    """Drop statistics of post/unionfind_periodic (CPU restatement)."""
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    n = phi.shape[0]
    out = np.zeros((max(1, max_drops), 10))
    nd = ctypes.c_int64()
    vel = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, w)]
    ptrs = [None if a is None else _dptr(a) for a in vel]
    rc = lib().oracle_drop_stats(n, _dptr(phi), *ptrs, float(threshold), int(connectivity),
                                 int(max_drops), _dptr(out), ctypes.byref(nd))
    if rc:
        raise RuntimeError(f"oracle_drop_stats failed: {rc}")
    k = min(nd.value, max_drops)
    cols = ["cells", "first_cell", "x", "y", "z", "u", "v", "w", "volume", "d_eq"]
    res = {c: out[:k, i].copy() for i, c in enumerate(cols)}
    res["count"] = nd.value
    return res


def fft3d(x: np.ndarray, sign: int = -1) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.complex128)
    n = x.shape[0]
    assert x.shape == (n, n, n)
    out = np.empty_like(x)
    lib().oracle_fft3d(n, x.ctypes.data_as(_dp), out.ctypes.data_as(_dp), sign)
    return out


class Oracle:
    """CPU FP64 restatement of the timestep with the C-ABI's shape."""

    def __init__(self, params: Params | None = None, **overrides):
        self.params = params if params is not None else default_params(**overrides)
        self.n = self.params.nx
        self._c = lib().oracle_create(ctypes.byref(self.params))
        if not self._c:
            raise RuntimeError("oracle_create failed")

    def close(self):
        if self._c:
            lib().oracle_destroy(self._c)
            self._c = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _field(self):
        return np.empty((self.n,) * 3, dtype=np.float64)

    def set_field(self, fieldn: int, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == (self.n,) * 3
        rc = lib().oracle_set_field(self._c, fieldn, _dptr(a))
        assert rc == 0, rc

    def get_field(self, fieldn: int) -> np.ndarray:
        out = self._field()
        rc = lib().oracle_get_field(self._c, fieldn, _dptr(out))
        assert rc == 0, rc
        return out

    def set_tracers(self, x, y, z):
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        lib().oracle_set_tracers(self._c, x.size, _dptr(x), _dptr(y), _dptr(z))
        self.np = x.size

    def get_tracers(self):
        cols = [np.empty(self.np) for _ in range(6)]
        lib().oracle_get_tracers(self._c, *[_dptr(a) for a in cols])
        return cols

    def set_ab2_state(self, alpha, beta, gumax):
        lib().oracle_set_ab2_state(self._c, alpha, beta, gumax)

    def step(self, nsteps: int = 1) -> int:
        return lib().oracle_step(self._c, int(nsteps))

    def diag(self) -> dict:
        d = Diag()
        lib().oracle_diag(self._c, ctypes.byref(d))
        return d.as_dict()

    def ns_rhs(self, parts: int = NS_CONV | NS_VISC | NS_FORCE):
        out = [self._field() for _ in range(3)]
        lib().oracle_ns_rhs(self._c, parts, *[_dptr(a) for a in out])
        return out

    def acdi_rhs(self, parts: int, gamma: float):
        out = self._field()
        rc = lib().oracle_acdi_rhs(self._c, parts, gamma, _dptr(out))
        assert rc == 0, rc
        return out

    def csf(self):
        out = [self._field() for _ in range(3)]
        rc = lib().oracle_csf(self._c, *[_dptr(a) for a in out])
        assert rc == 0, rc
        return out

    def divergence(self):
        out = self._field()
        lib().oracle_divergence(self._c, _dptr(out))
        return out

    def interp(self):
        out = [np.empty(self.np) for _ in range(3)]
        lib().oracle_interp(self._c, *[_dptr(a) for a in out])
        return out

    def poisson(self, rhs: np.ndarray) -> np.ndarray:
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        out = self._field()
        lib().oracle_poisson(self._c, _dptr(rhs), _dptr(out))
        return out
