"""ctypes mirror of include/mhit36.h (structs, constants, library loader).

The structs here must match the C declarations field for field; the
``test_abi`` CPU tests check the sizes against the compiled library.
"""
from __future__ import annotations

import ctypes
import os

OK = 0
ERR_ARG = 1
ERR_HIP = 2
ERR_CFL = 3
ERR_NOMEM = 4
ERR_STATE = 5
ERR_NOGPU = 6
ERR_IO = 7

POISSON_FP64 = 0
POISSON_FP32 = 1
K2_CONTINUOUS = 0
K2_MODIFIED = 1
INTERP_REFERENCE = 0
INTERP_TRILINEAR = 1

FIELD_U, FIELD_V, FIELD_W, FIELD_P, FIELD_PHI = 1, 2, 3, 4, 5
FIELD_RHSU_O, FIELD_RHSV_O, FIELD_RHSW_O, FIELD_RHSPHI_O = 6, 7, 8, 9
FIELD_NAMES = {FIELD_U: "u", FIELD_V: "v", FIELD_W: "w", FIELD_P: "p", FIELD_PHI: "phi",
               FIELD_RHSU_O: "rhsu_o", FIELD_RHSV_O: "rhsv_o", FIELD_RHSW_O: "rhsw_o",
               FIELD_RHSPHI_O: "rhsphi_o"}


class Params(ctypes.Structure):
    """mhit_params: input.inp (multi/input.inp:1-18) plus build switches."""

    _fields_ = [
        ("restart", ctypes.c_int32),
        ("tstart", ctypes.c_int32),
        ("tfin", ctypes.c_int32),
        ("dump", ctypes.c_int32),
        ("inflow", ctypes.c_int32),
        ("inphi", ctypes.c_int32),
        ("dt", ctypes.c_double),
        ("mu", ctypes.c_double),
        ("rho", ctypes.c_double),
        ("f1", ctypes.c_double),
        ("f2", ctypes.c_double),
        ("f3", ctypes.c_double),
        ("k0", ctypes.c_double),
        ("radius", ctypes.c_double),
        ("sigma", ctypes.c_double),
        ("epsr", ctypes.c_double),
        ("npart", ctypes.c_int64),
        ("inpart", ctypes.c_int32),
        ("nx", ctypes.c_int32),
        ("phiflag", ctypes.c_int32),
        ("partflag", ctypes.c_int32),
        ("poisson_precision", ctypes.c_int32),
        ("k2_mode", ctypes.c_int32),
        ("interp_mode", ctypes.c_int32),
        ("cfl_abort", ctypes.c_int32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Diag(ctypes.Structure):
    """mhit_diag: per-step scalars and diagnostics."""

    _fields_ = [
        ("step", ctypes.c_int64),
        ("alpha", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("gumax", ctypes.c_double),
        ("cou", ctypes.c_double),
        ("umean", ctypes.c_double),
        ("vmean", ctypes.c_double),
        ("wmean", ctypes.c_double),
        ("ke", ctypes.c_double),
        ("dissipation", ctypes.c_double),
        ("maxdiv", ctypes.c_double),
        ("phi_mass", ctypes.c_double),
        ("phi_min", ctypes.c_double),
        ("phi_max", ctypes.c_double),
        ("nan_count", ctypes.c_int64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Tuning(ctypes.Structure):
    """mhit_tuning: launch shapes (no result bit changes), the column
    transforms' form (fft_wave: rounding-level differences only) and the FFT
    launch widths (0 = the default; a width to or from eight columns moves
    a pass to or from the wavefront form), whether the x and y transforms
    of a plane share one launch (fft_xy) and the x passes' radix plan
    (fft_xplan); those two change rounding only."""

    _fields_ = [
        ("ns_mz", ctypes.c_int32),
        ("ac_mz", ctypes.c_int32),
        ("xcd", ctypes.c_int32),
        ("sort_every", ctypes.c_int32),
        ("ns_my", ctypes.c_int32),
        ("ac_my", ctypes.c_int32),
        ("fft_wave", ctypes.c_int32),
        ("fft_ycols", ctypes.c_int32),
        ("fft_zcols", ctypes.c_int32),
        ("fft_xrows", ctypes.c_int32),
        ("fft_xy", ctypes.c_int32),
        ("fft_xplan", ctypes.c_int32),
        ("ms_per_step", ctypes.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Drop(ctypes.Structure):
    """mhit_drop: one connected drop of the phase field."""

    _fields_ = [
        ("cells", ctypes.c_int64),
        ("first_cell", ctypes.c_int64),
        ("x", ctypes.c_double), ("y", ctypes.c_double), ("z", ctypes.c_double),
        ("u", ctypes.c_double), ("v", ctypes.c_double), ("w", ctypes.c_double),
        ("volume", ctypes.c_double),
        ("d_eq", ctypes.c_double),
    ]


PROF_MAX = 24


class Profile(ctypes.Structure):
    """mhit_profile: per-stage event timing of mhit_step."""

    _fields_ = [
        ("nstages", ctypes.c_int32),
        ("name", (ctypes.c_char * 32) * PROF_MAX),
        ("calls", ctypes.c_int64 * PROF_MAX),
        ("ms", ctypes.c_double * PROF_MAX),
        ("bytes_per_call", ctypes.c_double * PROF_MAX),
    ]

    def as_rows(self):
        rows = []
        for s in range(self.nstages):
            rows.append({
                "stage": bytes(self.name[s]).split(b"\0", 1)[0].decode(),
                "calls": int(self.calls[s]),
                "ms": float(self.ms[s]),
                "bytes_per_call": float(self.bytes_per_call[s]),
            })
        return rows


def default_params(**overrides) -> Params:
    """Reference defaults: multi/input.inp values with nx = 256
    (multi/module.f90:2), FP32 Poisson, continuous k^2."""
    p = Params(
        restart=0, tstart=0, tfin=10000, dump=100000, inflow=0, inphi=0,
        dt=0.0002, mu=0.006, rho=1.0, f1=1.0, f2=1.0, f3=1.0, k0=2.0,
        radius=0.5, sigma=1.0, epsr=1.0, npart=10000, inpart=1,
        nx=256, phiflag=0, partflag=0, poisson_precision=POISSON_FP32,
        k2_mode=K2_CONTINUOUS, interp_mode=INTERP_REFERENCE, cfl_abort=1,
    )
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise KeyError(f"unknown parameter {k!r}")
        setattr(p, k, v)
    return p


PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "libmhit36.so"
LIB_PATH = os.path.join(PKG_DIR, LIB_NAME)

# (name, restype, argtypes) for every entry point declared in include/mhit36.h
_P = ctypes.POINTER
_vp = ctypes.c_void_p
_dp = _P(ctypes.c_double)
EXPORTS = [
    ("mhit_params_default", ctypes.c_int, [_P(Params)]),
    ("mhit_create", ctypes.c_int, [_P(Params), _P(_vp)]),
    ("mhit_destroy", ctypes.c_int, [_vp]),
    ("mhit_upload_field", ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    ("mhit_download_field", ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    ("mhit_dump_async", ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_int32, ctypes.c_uint32]),
    ("mhit_dump_wait", ctypes.c_int, [_vp]),
    ("mhit_upload_tracers", ctypes.c_int, [_vp, ctypes.c_int64, _dp, _dp, _dp]),
    ("mhit_upload_tracer_state", ctypes.c_int, [_vp, ctypes.c_int64] + [_dp] * 7),
    ("mhit_download_tracers", ctypes.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    ("mhit_tracer_count", ctypes.c_int64, [_vp]),
    ("mhit_step", ctypes.c_int, [_vp, ctypes.c_int]),
    ("mhit_diag_compute", ctypes.c_int, [_vp, _P(Diag)]),
    ("mhit_scalars", ctypes.c_int, [_vp, _P(Diag)]),
    ("mhit_set_ab2_state", ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    ("mhit_sync", ctypes.c_int, [_vp]),
    ("mhit_get_tuning", ctypes.c_int, [_vp, _P(Tuning)]),
    ("mhit_set_tuning", ctypes.c_int, [_vp, _P(Tuning)]),
    ("mhit_autotune", ctypes.c_int, [_vp, ctypes.c_int, _P(Tuning)]),
    ("mhit_energy_spectrum", ctypes.c_int, [_vp, ctypes.c_int32, _dp, _P(ctypes.c_int32)]),
    ("mhit_energy_spectrum_noalloc", ctypes.c_int, [_vp, ctypes.c_int32, _dp, _P(ctypes.c_int32)]),
    ("mhit_drop_stats", ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int64, _P(Drop),
                                       _P(ctypes.c_int64)]),
    ("mhit_last_error", ctypes.c_char_p, []),
    ("mhit_init_taylor_green", ctypes.c_int, [_vp]),
    ("mhit_init_drop", ctypes.c_int, [_vp]),
    ("mhit_init_random_velocity", ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.c_double, ctypes.c_double]),
    ("mhit_init_drops", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_double, ctypes.c_uint64]),
    ("mhit_drop_centers", ctypes.c_int, [ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, _dp]),
    ("mhit_init_tracers_uniform", ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_uint64]),
    ("mhit_poisson_solve", ctypes.c_int, [_vp, _dp, _dp]),
    ("mhit_fft_r2c", ctypes.c_int, [_vp, _dp, _dp]),
    ("mhit_fft_c2r", ctypes.c_int, [_vp, _dp, _dp]),
    ("mhit_profile_enable", ctypes.c_int, [_vp, ctypes.c_int]),
    ("mhit_profile_read", ctypes.c_int, [_vp, _P(Profile)]),
    ("mhit_copy_probe", ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int, _dp]),
    ("mhit_stream_probe", ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]),
    ("mhit_ensemble_create", ctypes.c_int, [ctypes.c_int32, _P(Params), _P(_vp)]),
    ("mhit_ensemble_box", ctypes.c_int, [_vp, ctypes.c_int32, _P(_vp)]),
    ("mhit_ensemble_step", ctypes.c_int, [_vp, ctypes.c_int]),
    ("mhit_ensemble_set_batched", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("mhit_ensemble_autotune", ctypes.c_int, [_vp, ctypes.c_int, _P(ctypes.c_double)]),
    ("mhit_ensemble_destroy", ctypes.c_int, [_vp]),
    ("mhit_version", ctypes.c_char_p, []),
]

_lib = None


def load_library(path: str | None = None) -> ctypes.CDLL:
    """Load libmhit36.so (built in-tree by __graft_entry__.build()).

    Raises if the library is missing: there is no fallback path."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} not found: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(path)
    for name, res, args in EXPORTS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path == LIB_PATH:
        _lib = lib
    return lib
