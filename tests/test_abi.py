"""The C-ABI library loads on a machine without a GPU, exports every entry
point declared in include/mhit36.h, lays out its structs as the ctypes
mirror does, and fails loudly (no CPU fallback) when no device is there."""
import ctypes
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "mhit36.h")


@pytest.fixture(scope="module")
def lib():
    from mhit36_mixedp_amd import _abi
    if not os.path.exists(_abi.LIB_PATH):
        from mhit36_mixedp_amd import build
        build.build(verbose=False)
    return _abi.load_library()


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mhit_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(lib):
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    from mhit36_mixedp_amd import _abi
    assert sorted(n for n, _, _ in _abi.EXPORTS) == names


def test_step_stages_carry_roctx_ranges(lib):
    """The timestep and its stages are wrapped in roctx ranges (the
    reference's NVTX ranges, main.f90:460-1040), so the library imports
    the roctx push/pop pair and names the reference's range labels."""
    path = os.path.join(ROOT, "mhit36_mixedp_amd", "libmhit36.so")
    und = subprocess.run(["nm", "-D", "--undefined-only", path], capture_output=True, text=True,
                         check=True).stdout
    assert "roctxRangePushA" in und and "roctxRangePop" in und
    with open(path, "rb") as f:
        blob = f.read()
    for label in (b"Iteration", b"Phase-field", b"Projection", b"Poisson", b"Correction"):
        assert label + b"\0" in blob, label


def test_struct_layout_matches_c():
    from mhit36_mixedp_amd import _abi
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "mhit36.h"
#define F(T, m) printf(#T "." #m " %zu\n", offsetof(T, m));
int main(void) {
  printf("mhit_params %zu\n", sizeof(mhit_params));
  printf("mhit_diag %zu\n", sizeof(mhit_diag));
  printf("mhit_profile %zu\n", sizeof(mhit_profile));
  printf("mhit_tuning %zu\n", sizeof(mhit_tuning));
  F(mhit_tuning, fft_wave) F(mhit_tuning, fft_xrows) F(mhit_tuning, fft_xy) F(mhit_tuning, fft_xplan) F(mhit_tuning, ms_per_step)
  F(mhit_params, dt) F(mhit_params, npart) F(mhit_params, nx) F(mhit_params, cfl_abort)
  F(mhit_diag, ke) F(mhit_diag, nan_count) F(mhit_profile, ms) F(mhit_profile, bytes_per_call)
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = dict(l.rsplit(" ", 1) for l in subprocess.run([exe], capture_output=True, text=True,
                                                             check=True).stdout.splitlines())
    assert int(out["mhit_params"]) == ctypes.sizeof(_abi.Params)
    assert int(out["mhit_diag"]) == ctypes.sizeof(_abi.Diag)
    assert int(out["mhit_profile"]) == ctypes.sizeof(_abi.Profile)
    assert int(out["mhit_tuning"]) == ctypes.sizeof(_abi.Tuning)
    for key, val in out.items():
        if "." in key:
            t, m = key.split(".")
            cls = {"mhit_params": _abi.Params, "mhit_diag": _abi.Diag, "mhit_profile": _abi.Profile,
                   "mhit_tuning": _abi.Tuning}[t]
            assert getattr(cls, m).offset == int(val), key


def test_defaults_are_input_inp(lib):
    """multi/input.inp values (the driver's input.inp holds the same) and
    nx = 256 (multi/module.f90:2)."""
    from mhit36_mixedp_amd import _abi
    p = _abi.Params()
    assert lib.mhit_params_default(ctypes.byref(p)) == 0
    assert (p.restart, p.tstart, p.tfin, p.dump, p.inflow, p.inphi) == (0, 0, 10000, 100000, 0, 0)
    assert (p.dt, p.mu, p.rho, p.f1, p.f2, p.f3, p.k0) == (0.0002, 0.006, 1.0, 1.0, 1.0, 1.0, 2.0)
    assert (p.radius, p.sigma, p.epsr, p.npart, p.inpart) == (0.5, 1.0, 1.0, 10000, 1)
    assert (p.nx, p.phiflag, p.partflag, p.poisson_precision) == (256, 0, 0, _abi.POISSON_FP32)
    q = _abi.default_params()
    assert q.as_dict() == p.as_dict()


def test_bad_size_is_rejected_before_any_device_call(lib):
    from mhit36_mixedp_amd import _abi
    p = _abi.default_params(nx=100)
    h = ctypes.c_void_p()
    assert lib.mhit_create(ctypes.byref(p), ctypes.byref(h)) == _abi.ERR_ARG
    assert b"unsupported" in lib.mhit_last_error()
    assert not h.value


def test_no_device_fails_loudly(lib):
    """Without a HIP device mhit_create returns MHIT_ERR_NOGPU: there is no
    CPU fallback.  (On a GPU machine it succeeds instead.)"""
    from mhit36_mixedp_amd import _abi
    p = _abi.default_params(nx=16)
    h = ctypes.c_void_p()
    rc = lib.mhit_create(ctypes.byref(p), ctypes.byref(h))
    if rc == _abi.OK:
        lib.mhit_destroy(h)
        pytest.skip("a HIP device is present")
    assert rc == _abi.ERR_NOGPU
    assert b"no HIP device" in lib.mhit_last_error()
    from mhit36_mixedp_amd import MhitError, Solver
    with pytest.raises(MhitError):
        Solver(nx=16)


def test_drop_centers_are_separated_and_seeded(lib):
    from mhit36_mixedp_amd import drop_centers
    c1 = drop_centers(64, 0.2, 0.1, 7)
    c2 = drop_centers(64, 0.2, 0.1, 7)
    c3 = drop_centers(64, 0.2, 0.1, 8)
    assert np.array_equal(c1, c2) and not np.array_equal(c1, c3)
    L = 2 * np.pi
    assert c1.min() >= 0 and c1.max() < L
    d = np.abs(c1[:, None, :] - c1[None, :, :])
    d = np.minimum(d, L - d)
    r = np.sqrt((d ** 2).sum(-1)) + np.eye(64) * 10
    assert r.min() >= 2 * 0.2 + 0.1
