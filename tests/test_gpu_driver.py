"""The Fortran driver (driver/, iso_c_binding over the C-ABI) runs the
reference's program flow: input.inp, fresh start with Taylor-Green + a
centred drop + tracers, dumps to ./output/<f>_%08d.dat, and a restart from
those files.  Its outputs equal the Python host's on the same inputs."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "driver", "mhit36")


def write_inp(path, restart, tstart, tfin, dump, inflow=0, inphi=0, npart=3000):
    vals = [restart, tstart, tfin, dump, inflow, inphi, 0.001, 0.01, 1.0, 1.0, 1.0, 1.0, 2.0,
            1.2, 0.5, 1.0, npart, 1]
    with open(os.path.join(path, "input.inp"), "w") as f:
        for v in vals:
            f.write(f"{v}      ! value\n")


def read(path, name, n):
    a = np.fromfile(os.path.join(path, "output", name), dtype=np.float64)
    return a.reshape(n, n, n)


def run(path, *args):
    r = subprocess.run([DRIVER, *args], cwd=path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def test_driver_fresh_start_and_restart(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.skip("driver not built (no amdflang)")
    from mhit36_mixedp_amd import Solver
    n = 16
    os.makedirs(tmp_path / "output")
    write_inp(tmp_path, 0, 0, 6, 3)
    out = run(str(tmp_path), "--nx", str(n), "--phi", "1", "--part", "1", "--seed", "5")
    assert re.search(r"Time step\s+6\s+of\s+6", out), out[-500:]
    for t in (0, 3, 6):
        for f in ("u", "v", "w", "p", "phi"):
            assert os.path.exists(tmp_path / "output" / f"{f}_{t:08d}.dat")
    kw = dict(nx=n, phiflag=1, partflag=1, dt=0.001, mu=0.01, radius=1.2, sigma=0.5, npart=3000)
    with Solver(**kw) as s:
        s.init_taylor_green()
        s.init_drop()
        s.init_tracers_uniform(3000, 5)
        assert np.array_equal(s.download_field(1), read(tmp_path, "u_00000000.dat", n))
        s.step(6)
        for f, name in ((1, "u"), (2, "v"), (3, "w"), (4, "p"), (5, "phi")):
            assert np.array_equal(s.download_field(f), read(tmp_path, f"{name}_00000006.dat", n)), name
        t = s.download_tracers()
        raw = np.fromfile(tmp_path / "output" / "part_00000006.dat").reshape(4, 3000)
        assert np.array_equal(raw[1], t["x"]) and np.array_equal(raw[3], t["z"])
    # restart from the dump at t = 3: Euler first step (main.f90:440-441)
    write_inp(tmp_path, 1, 3, 5, 5)
    run(str(tmp_path), "--nx", str(n), "--phi", "1", "--quiet")
    with Solver(**kw) as s:
        for f, name in ((1, "u"), (2, "v"), (3, "w"), (5, "phi")):
            s.upload_field(f, read(tmp_path, f"{name}_00000003.dat", n))
        s.step(2)
        assert np.array_equal(s.download_field(1), read(tmp_path, "u_00000005.dat", n))
        assert np.array_equal(s.download_field(5), read(tmp_path, "phi_00000005.dat", n))


def test_driver_drop_centres(tmp_path, oracle_mod):
    """--drops writes post/unionfind_periodic's center_timestep file at each
    dump; its rows equal the oracle's statistics of the dumped fields."""
    if not os.path.exists(DRIVER):
        pytest.skip("driver not built (no amdflang)")
    n = 32
    os.makedirs(tmp_path / "output")
    write_inp(tmp_path, 0, 0, 4, 2, npart=0)
    run(str(tmp_path), "--nx", str(n), "--phi", "1", "--drops", "--quiet")
    phi = read(tmp_path, "phi_00000004.dat", n)
    u, v, w = (read(tmp_path, f"{f}_00000004.dat", n) for f in ("u", "v", "w"))
    o = oracle_mod.drop_stats(phi, u, v, w)
    rows = np.loadtxt(tmp_path / "output" / "center_timestep_00000004.dat", comments="#", ndmin=2)
    assert rows.shape[0] == o["count"] >= 1
    for col, k in enumerate(("x", "y", "z", "u", "v", "w", "volume"), start=1):
        assert np.allclose(rows[:, col], o[k], rtol=1e-8, atol=1e-10), k
    counts = np.loadtxt(tmp_path / "output" / "drop_count.dat", ndmin=2)
    assert counts[:, 0].tolist() == [0, 2, 4] and counts[-1, 1] == o["count"]
