"""Python host mirror of the C-ABI (include/mhit36.h) over ctypes.

``Solver`` plays the role of the reference's ``program main``
(multi/main.f90): it owns one device context, takes input.inp parameters,
loads initial fields, advances the timestep and reads fields back.  Every
call goes to the HIP library; there is no CPU path.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _abi
from ._abi import Diag, Params, default_params

_dp = ctypes.POINTER(ctypes.c_double)


class MhitError(RuntimeError):
    def __init__(self, code: int, where: str):
        lib = _abi.load_library()
        msg = lib.mhit_last_error().decode(errors="replace")
        super().__init__(f"{where} failed with status {code}: {msg}")
        self.code = code


class CFLStop(MhitError):
    """cou > 7: the reference stops the run (multi/main.f90:1006)."""


def _check(rc: int, where: str):
    if rc == _abi.OK:
        return
    if rc == _abi.ERR_CFL:
        raise CFLStop(rc, where)
    raise MhitError(rc, where)


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def version() -> str:
    return _abi.load_library().mhit_version().decode()


def drop_centers(ndrops: int, radius: float, min_gap: float, seed: int) -> np.ndarray:
    """Centres placed by ``Solver.init_drops`` (host-side generator)."""
    out = np.zeros((max(ndrops, 1), 3))
    _check(_abi.load_library().mhit_drop_centers(ndrops, radius, min_gap, seed, _ptr(out)),
           "mhit_drop_centers")
    return out[:ndrops]


class Solver:
    """One MHIT36 box on one GPU."""

    def __init__(self, params: Params | None = None, _handle=None, **overrides):
        self.lib = _abi.load_library()
        self.params = params if params is not None else default_params(**overrides)
        if params is not None and overrides:
            for k, v in overrides.items():
                setattr(self.params, k, v)
        self.n = int(self.params.nx)
        self._owner = _handle is None
        if _handle is not None:          # a box owned by an Ensemble
            self._h = _handle
            return
        h = ctypes.c_void_p()
        _check(self.lib.mhit_create(ctypes.byref(self.params), ctypes.byref(h)), "mhit_create")
        self._h = h

    # context management ---------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            if self._owner:
                self.lib.mhit_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # fields ---------------------------------------------------------------
    def upload_field(self, fieldn: int, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != (self.n,) * 3:
            raise ValueError(f"field must be {(self.n,) * 3} (z, y, x), got {a.shape}")
        _check(self.lib.mhit_upload_field(self._h, fieldn, _ptr(a)), "mhit_upload_field")

    def download_field(self, fieldn: int) -> np.ndarray:
        out = np.empty((self.n,) * 3, dtype=np.float64)
        _check(self.lib.mhit_download_field(self._h, fieldn, _ptr(out)), "mhit_download_field")
        return out

    def dump_async(self, directory: str, step: int, fields=(1, 2, 3, 4)):
        """Write the fields as <directory>/<name>_%08d.dat (the reference's
        output files) from a host thread; returns at once.  A second call
        waits for the first; ``dump_wait`` waits for the pending one."""
        mask = 0
        for f in fields:
            mask |= 1 << int(f)
        _check(self.lib.mhit_dump_async(self._h, str(directory).encode(), int(step), mask),
               "mhit_dump_async")

    def dump_wait(self):
        _check(self.lib.mhit_dump_wait(self._h), "mhit_dump_wait")

    def set_velocity(self, u, v, w):
        for f, a in zip((_abi.FIELD_U, _abi.FIELD_V, _abi.FIELD_W), (u, v, w)):
            self.upload_field(f, a)

    def velocity(self):
        return tuple(self.download_field(f) for f in (_abi.FIELD_U, _abi.FIELD_V, _abi.FIELD_W))

    # tracers ----------------------------------------------------------------
    def upload_tracers(self, x, y, z):
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        _check(self.lib.mhit_upload_tracers(self._h, x.size, _ptr(x), _ptr(y), _ptr(z)),
               "mhit_upload_tracers")

    def download_tracers(self) -> dict:
        n = self.lib.mhit_tracer_count(self._h)
        cols = {k: np.empty(n) for k in ("tag", "x", "y", "z", "up", "vp", "wp")}
        _check(self.lib.mhit_download_tracers(self._h, *[_ptr(cols[k]) for k in
                                                         ("tag", "x", "y", "z", "up", "vp", "wp")]),
               "mhit_download_tracers")
        return cols

    def upload_tracer_state(self, cols: dict):
        """Full tracer state as returned by ``download_tracers``."""
        keys = ("tag", "x", "y", "z", "up", "vp", "wp")
        a = [np.ascontiguousarray(cols[k], dtype=np.float64) for k in keys]
        _check(self.lib.mhit_upload_tracer_state(self._h, a[0].size, *[_ptr(x) for x in a]),
               "mhit_upload_tracer_state")

    # checkpoint ---------------------------------------------------------------
    def checkpoint(self) -> dict:
        """Everything the next step reads: velocity, phase field, the AB2
        history (rhs*_o, tracer velocities) and alpha/beta/gumax.  Restoring
        it into a fresh context continues the run bit for bit; the reference
        restarts with an Euler step instead (multi/main.f90:440-441)."""
        fields = [_abi.FIELD_U, _abi.FIELD_V, _abi.FIELD_W,
                  _abi.FIELD_RHSU_O, _abi.FIELD_RHSV_O, _abi.FIELD_RHSW_O]
        if self.params.phiflag:
            fields += [_abi.FIELD_PHI, _abi.FIELD_RHSPHI_O]
        d = self.diag()
        state = {"fields": {f: self.download_field(f) for f in fields},
                 "alpha": d["alpha"], "beta": d["beta"], "gumax": d["gumax"]}
        if self.lib.mhit_tracer_count(self._h) > 0:
            state["tracers"] = self.download_tracers()
        return state

    def restore(self, state: dict):
        for f, a in state["fields"].items():
            self.upload_field(f, a)
        self.set_ab2_state(state["alpha"], state["beta"], state["gumax"])
        if "tracers" in state:
            self.upload_tracer_state(state["tracers"])

    # time stepping ----------------------------------------------------------
    def step(self, nsteps: int = 1):
        _check(self.lib.mhit_step(self._h, int(nsteps)), "mhit_step")

    def set_ab2_state(self, alpha: float, beta: float, gumax: float):
        _check(self.lib.mhit_set_ab2_state(self._h, alpha, beta, gumax), "mhit_set_ab2_state")

    def sync(self):
        _check(self.lib.mhit_sync(self._h), "mhit_sync")

    def diag(self) -> dict:
        d = Diag()
        _check(self.lib.mhit_diag_compute(self._h, ctypes.byref(d)), "mhit_diag_compute")
        return d.as_dict()

    def scalars(self) -> dict:
        """The last step's scalars (step, alpha, beta, gumax, cou, means)
        without a pass over the fields."""
        d = Diag()
        _check(self.lib.mhit_scalars(self._h, ctypes.byref(d)), "mhit_scalars")
        return d.as_dict()

    def tuning(self) -> dict:
        t = _abi.Tuning()
        _check(self.lib.mhit_get_tuning(self._h, ctypes.byref(t)), "mhit_get_tuning")
        return t.as_dict()

    def set_tuning(self, **kw):
        t = _abi.Tuning()
        _check(self.lib.mhit_get_tuning(self._h, ctypes.byref(t)), "mhit_get_tuning")
        for k, v in kw.items():
            setattr(t, k, v)
        _check(self.lib.mhit_set_tuning(self._h, ctypes.byref(t)), "mhit_set_tuning")

    def autotune(self, steps: int = 5) -> dict:
        """Pick the fastest launch shapes on the current state; the state is
        restored afterwards."""
        t = _abi.Tuning()
        _check(self.lib.mhit_autotune(self._h, int(steps), ctypes.byref(t)), "mhit_autotune")
        return t.as_dict()

    def energy_spectrum(self, noalloc: bool = False) -> np.ndarray:
        """E(k), k = 0, 1, ...: shell spectrum of the kinetic energy.
        noalloc=True uses mhit_energy_spectrum_noalloc (no device
        allocation; shell sums in the library's fixed per-device array)."""
        nb = int(np.floor(np.sqrt(3.0) * (self.n // 2) + 0.5)) + 1
        E = np.zeros(nb)
        ns = ctypes.c_int32()
        fn = "mhit_energy_spectrum_noalloc" if noalloc else "mhit_energy_spectrum"
        _check(getattr(self.lib, fn)(self._h, nb, _ptr(E), ctypes.byref(ns)), fn)
        return E[:ns.value]

    def drop_stats(self, threshold: float = 0.5, connectivity: int = 6, max_drops: int = 1 << 16):
        """Connected drops of phi >= threshold (see mhit_drop_stats); a dict of
        numpy columns, one entry per drop, ordered by first cell."""
        buf = (_abi.Drop * max(1, max_drops))()
        nd = ctypes.c_int64()
        _check(self.lib.mhit_drop_stats(self._h, float(threshold), int(connectivity), int(max_drops),
                                        buf, ctypes.byref(nd)), "mhit_drop_stats")
        k = min(nd.value, max_drops)
        names = [f for f, _ in _abi.Drop._fields_]
        out = {f: np.array([getattr(buf[i], f) for i in range(k)]) for f in names}
        out["count"] = nd.value
        return out

    def profile_enable(self, on: bool = True):
        _check(self.lib.mhit_profile_enable(self._h, int(on)), "mhit_profile_enable")

    def copy_probe(self, nbytes: int = 1 << 30, iters: int = 10) -> float:
        """Plain device copy rate in GB/s (bytes read + written)."""
        g = ctypes.c_double()
        _check(self.lib.mhit_copy_probe(self._h, int(nbytes), int(iters), ctypes.byref(g)),
               "mhit_copy_probe")
        return g.value

    def stream_probe(self, nbytes_per_stream: int = 512 << 20, reads: int = 7, writes: int = 6,
                     iters: int = 5) -> float:
        """Rate in GB/s (read + write) of a pointwise kernel over `reads`
        input and `writes` output FP64 arrays: the ceiling for a pass with
        that stream mix."""
        g = ctypes.c_double()
        _check(self.lib.mhit_stream_probe(self._h, int(nbytes_per_stream), int(reads), int(writes),
                                          int(iters), ctypes.byref(g)), "mhit_stream_probe")
        return g.value

    def profile(self) -> list:
        """Per-stage event timing accumulated since profile_enable()."""
        p = _abi.Profile()
        _check(self.lib.mhit_profile_read(self._h, ctypes.byref(p)), "mhit_profile_read")
        return [r for r in p.as_rows() if r["calls"] > 0]

    # initial conditions -------------------------------------------------------
    def init_taylor_green(self):
        _check(self.lib.mhit_init_taylor_green(self._h), "mhit_init_taylor_green")

    def init_drop(self):
        _check(self.lib.mhit_init_drop(self._h), "mhit_init_drop")

    def init_random_velocity(self, seed: int, kmax: float = 8.0, urms: float = 1.0):
        _check(self.lib.mhit_init_random_velocity(self._h, seed, kmax, urms),
               "mhit_init_random_velocity")

    def init_drops(self, ndrops: int, radius: float, seed: int):
        _check(self.lib.mhit_init_drops(self._h, ndrops, radius, seed), "mhit_init_drops")

    def init_tracers_uniform(self, n: int, seed: int):
        _check(self.lib.mhit_init_tracers_uniform(self._h, n, seed), "mhit_init_tracers_uniform")

    # building blocks ------------------------------------------------------------
    def poisson_solve(self, rhs: np.ndarray) -> np.ndarray:
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        out = np.empty_like(rhs)
        _check(self.lib.mhit_poisson_solve(self._h, _ptr(rhs), _ptr(out)), "mhit_poisson_solve")
        return out

    def fft_r2c(self, x: np.ndarray) -> np.ndarray:
        """Unnormalised forward 3-D R2C FFT; returns complex [kz, ky, kx<=n/2]."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        n = self.n
        out = np.empty((n, n, n // 2 + 1, 2))
        _check(self.lib.mhit_fft_r2c(self._h, _ptr(x), _ptr(out)), "mhit_fft_r2c")
        return out[..., 0] + 1j * out[..., 1]

    def fft_c2r(self, X: np.ndarray) -> np.ndarray:
        """Unnormalised inverse 3-D C2R FFT of a [kz, ky, kx<=n/2] spectrum."""
        n = self.n
        X = np.ascontiguousarray(X, dtype=np.complex128)
        assert X.shape == (n, n, n // 2 + 1)
        buf = np.ascontiguousarray(np.stack([X.real, X.imag], axis=-1))
        out = np.empty((n, n, n))
        _check(self.lib.mhit_fft_c2r(self._h, _ptr(buf), _ptr(out)), "mhit_fft_c2r")
        return out


class Ensemble:
    """Independent boxes advanced together on one GPU (mhit_ensemble_*).

    ``boxes[i]`` is a Solver view of box i (initialise, upload, download,
    diagnose through it); ``step(n)`` advances every box n steps as one
    captured graph per step.  ``batched`` (default True): boxes of one grid
    size share every launch of a pass; False: one graph branch per box.
    The results are the same bits either way."""

    def __init__(self, params_list, batched: bool = True):
        self.lib = _abi.load_library()
        arr = (Params * len(params_list))(*params_list)
        h = ctypes.c_void_p()
        _check(self.lib.mhit_ensemble_create(len(params_list), arr, ctypes.byref(h)),
               "mhit_ensemble_create")
        self._h = h
        self.boxes = []
        for i, p in enumerate(params_list):
            b = ctypes.c_void_p()
            _check(self.lib.mhit_ensemble_box(h, i, ctypes.byref(b)), "mhit_ensemble_box")
            self.boxes.append(Solver(p, _handle=b))
        if not batched:
            self.set_batched(False)

    def set_batched(self, on: bool):
        _check(self.lib.mhit_ensemble_set_batched(self._h, int(bool(on))), "mhit_ensemble_set_batched")

    def step(self, nsteps: int = 1):
        _check(self.lib.mhit_ensemble_step(self._h, int(nsteps)), "mhit_ensemble_step")

    def autotune(self, steps: int = 3) -> float:
        """Tune the batched launch shapes per grid size on the current
        state (restored afterwards); returns ms per ensemble step with the
        chosen settings (each box's ``tuning()`` reports them)."""
        ms = ctypes.c_double()
        _check(self.lib.mhit_ensemble_autotune(self._h, int(steps), ctypes.byref(ms)),
               "mhit_ensemble_autotune")
        return ms.value

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            for b in self.boxes:
                b.close()
            self.lib.mhit_ensemble_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
