"""mhit_energy_spectrum_noalloc (no device allocation; shell sums in the
library's fixed per-device array) against mhit_energy_spectrum, through the
C-ABI: two calls in a row, the second on a smaller box, so the second call
finds the first call's sums in the array and must clear them.  The shell
sums are atomic adds in both entry points, so they agree to rounding:
1e-12 relative to the total energy in FP64, as the spectrum's own bar."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _spectra(s, nb):
    out = []
    for fn in ("mhit_energy_spectrum", "mhit_energy_spectrum_noalloc"):
        E = np.full(nb, np.nan)
        ns = ctypes.c_int32(-1)
        rc = getattr(s.lib, fn)(s._h, nb, E.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                ctypes.byref(ns))
        assert rc == 0, (fn, s.lib.mhit_last_error())
        out.append((E, ns.value))
    return out


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
def test_noalloc_spectrum_matches_allocating_entry_point(prec):
    from mhit36_mixedp_amd import Solver
    for n, seed in ((64, 5), (24, 9)):           # the second call on a smaller box
        print(f"[noalloc spectrum] n={n} prec={prec}", flush=True)
        with Solver(nx=n, poisson_precision=prec, dt=5e-4) as s:
            s.init_random_velocity(seed, kmax=6.0, urms=1.0)
            s.step(2)
            nb = int(np.floor(np.sqrt(3.0) * (n // 2) + 0.5)) + 1
            (Ea, na), (Eb, nbs) = _spectra(s, nb + 3)   # 3 bins past the last shell
            assert na == nbs == nb
            assert np.all(Eb[nb:] == 0.0) and np.all(Ea[nb:] == 0.0)
            tot = Ea.sum()
            assert tot > 0
            assert np.max(np.abs(Eb - Ea)) < 1e-12 * tot
            assert abs(Eb.sum() - s.diag()["ke"]) < (1e-12 if prec == 0 else 1e-6) * tot


def test_noalloc_spectrum_rejects_too_many_shells():
    """Asking for fewer bins than the limit always works; the limit itself
    is only reachable for n > 4700, so the error path is checked through
    the argument checks shared with the allocating entry point."""
    from mhit36_mixedp_amd import Solver
    with Solver(nx=16, poisson_precision=0) as s:
        E = np.zeros(4)
        ns = ctypes.c_int32()
        rc = s.lib.mhit_energy_spectrum_noalloc(s._h, 0, E.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                ctypes.byref(ns))
        assert rc != 0
        rc = s.lib.mhit_energy_spectrum_noalloc(s._h, 4, E.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                ctypes.byref(ns))
        assert rc == 0 and ns.value == 15
