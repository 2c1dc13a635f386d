"""A checkpoint restored into a fresh context continues the run bit for bit.

The reference restarts from u, v, w, phi dumps with an Euler step
(multi/main.f90:440-441), so its restarted run departs from the
uninterrupted one.  Persisting the AB2 history (rhs*_o, rhsphi_o, the
tracers' previous-step velocities, alpha/beta/gumax) removes that
departure: checked through the Python host and through the Fortran driver's
``--exact-restart`` dumps."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def _by_tag(t):
    o = np.argsort(t["tag"])
    return {k: v[o] for k, v in t.items()}


@pytest.mark.parametrize("fp", [0, 1])
def test_checkpoint_restore_bitwise(fp):
    from mhit36_mixedp_amd import Solver
    print(f"[restart] 32^3 multiphase + tracers, poisson_precision={fp}", flush=True)
    kw = dict(nx=32, phiflag=1, partflag=1, poisson_precision=fp, dt=5e-4, mu=0.006, sigma=0.5)
    with Solver(**kw) as a, Solver(**kw) as b:
        a.init_random_velocity(3, kmax=4.0, urms=1.0)
        a.init_drops(3, 0.6, 3)
        a.init_tracers_uniform(2500, 3)
        a.step(4)
        b.restore(a.checkpoint())
        a.step(5)
        b.step(5)
        for f in (1, 2, 3, 4, 5, 6, 7, 8, 9):
            assert np.array_equal(a.download_field(f), b.download_field(f)), f
        ta, tb = _by_tag(a.download_tracers()), _by_tag(b.download_tracers())
        for k in ta:
            assert np.array_equal(ta[k], tb[k]), k
        da, db = a.diag(), b.diag()
        da.pop("step"), db.pop("step")
        assert da == db


def test_driver_exact_restart(tmp_path):
    drv = os.path.join(ROOT, "driver", "mhit36")
    if not os.path.exists(drv):
        pytest.skip("driver not built (no amdflang)")
    from test_gpu_driver import read, write_inp
    print("[restart] driver --exact-restart 16^3", flush=True)
    n = 16
    args = ["--nx", str(n), "--phi", "1", "--part", "1", "--seed", "4", "--exact-restart", "--quiet"]
    os.makedirs(tmp_path / "output")
    write_inp(tmp_path, 0, 0, 6, 3)
    r = subprocess.run([drv, *args], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    for f in ("u", "phi", "rhsu_o", "rhsphi_o", "part", "partab2", "ab2"):
        assert os.path.exists(tmp_path / "output" / f"{f}_00000003.dat"), f
    keep = {f: read(tmp_path, f"{f}_00000006.dat", n) for f in ("u", "v", "w", "p", "phi")}
    part6 = np.fromfile(tmp_path / "output" / "part_00000006.dat")
    for f in keep:
        os.remove(tmp_path / "output" / f"{f}_00000006.dat")
    write_inp(tmp_path, 1, 3, 6, 3)
    r = subprocess.run([drv, *args], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    for f, a in keep.items():
        assert np.array_equal(read(tmp_path, f"{f}_00000006.dat", n), a), f
    p6 = np.fromfile(tmp_path / "output" / "part_00000006.dat")
    a, b = part6.reshape(4, -1), p6.reshape(4, -1)
    assert np.array_equal(a[:, np.argsort(a[0])], b[:, np.argsort(b[0])])
