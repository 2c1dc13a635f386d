"""MI355X-native single-GPU MHIT36 time-stepper (HIP, gfx950).

The product is ``libmhit36.so`` (C-ABI in include/mhit36.h); this package
holds its sources (csrc/), its build (build.py) and a ctypes host mirror
(api.py).
"""
from ._abi import (  # noqa: F401
    FIELD_P, FIELD_PHI, FIELD_U, FIELD_V, FIELD_W, INTERP_REFERENCE, INTERP_TRILINEAR,
    K2_CONTINUOUS, K2_MODIFIED, POISSON_FP32, POISSON_FP64, Diag, Params, default_params,
    load_library,
)
from .api import CFLStop, Ensemble, MhitError, Solver, drop_centers, version  # noqa: F401
