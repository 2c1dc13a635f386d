"""Asynchronous field dumps (mhit_dump_async / mhit_dump_wait).

The files must hold exactly what mhit_download_field returns at the time
of the call (the state is snapshot in stream order, the following steps
run while the host thread writes), named as the reference's writefield
names them (multi/readwrite.f90:38-98: output/<name>_%08d.dat, FP64,
Fortran order).  File-system errors come back as MHIT_ERR_IO."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NAMES = {1: "u", 2: "v", 3: "w", 4: "p", 5: "phi", 6: "rhsu_o", 7: "rhsv_o", 8: "rhsw_o", 9: "rhsphi_o"}


def _solver(prec, phase=True, n=32):
    from mhit36_mixedp_amd import Solver, default_params
    s = Solver(default_params(nx=n, phiflag=int(phase), partflag=0, poisson_precision=prec,
                              dt=5e-4, mu=0.006, sigma=0.5))
    s.init_random_velocity(7, kmax=4.0, urms=1.0)
    if phase:
        s.init_drops(2, 0.6, 7)
    return s


@pytest.mark.parametrize("prec", [0, 1])
def test_dump_files_equal_download_at_call(tmp_path, prec):
    print(f"[dump] 32^3 multiphase, poisson_precision={prec}", flush=True)
    s = _solver(prec)
    s.step(3)
    fields = list(range(1, 10))
    want = {f: s.download_field(f) for f in fields}
    s.dump_async(str(tmp_path), 3, fields)
    s.step(4)                      # runs while the files are written
    s.dump_wait()
    for f in fields:
        path = tmp_path / f"{NAMES[f]}_00000003.dat"
        got = np.fromfile(path, dtype=np.float64)
        assert got.size == 32 ** 3, f
        assert np.array_equal(got.reshape(32, 32, 32), want[f]), f
    # a second dump at a later step, fewer fields: the same buffers reused
    want2 = {f: s.download_field(f) for f in (1, 4)}
    s.dump_async(str(tmp_path), 7, (1, 4))
    s.dump_wait()
    for f in (1, 4):
        got = np.fromfile(tmp_path / f"{NAMES[f]}_00000007.dat", dtype=np.float64)
        assert np.array_equal(got.reshape(32, 32, 32), want2[f]), f
    s.close()


def test_dump_errors(tmp_path):
    from mhit36_mixedp_amd import MhitError
    from mhit36_mixedp_amd import _abi
    print("[dump] error paths", flush=True)
    s = _solver(1, phase=False, n=16)
    s.step(1)
    with pytest.raises(MhitError) as e:
        s.dump_async(str(tmp_path), 1, (5,))          # phi without a phase field
    assert e.value.code == _abi.ERR_ARG
    missing = os.path.join(str(tmp_path), "no", "such", "dir")
    s.dump_async(missing, 1, (1, 2))                 # starts; the writer fails
    with pytest.raises(MhitError) as e:
        s.dump_wait()
    assert e.value.code == _abi.ERR_IO
    s.dump_wait()                                     # the error was reported once
    # the error of a dump not waited for is returned by the next call
    s.dump_async(missing, 2, (1,))
    with pytest.raises(MhitError) as e:
        s.dump_async(str(tmp_path), 2, (1,))
    assert e.value.code == _abi.ERR_IO
    s.dump_async(str(tmp_path), 2, (1,))
    s.dump_wait()
    assert (tmp_path / "u_00000002.dat").stat().st_size == 16 ** 3 * 8
    s.close()
