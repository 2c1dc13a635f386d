"""FP32-Poisson form against the FP64-Poisson form on the GPU.

Both forms start from the same seeded fields (velocity, drops, tracers,
generated on the device) and advance the same number of steps; the FP64
form is the reference's own FP64 path (multi/main.f90:853-913 with
double-complex cuFFT) and serves as the FP64 oracle.  Reported: the relative
error of kinetic energy, dissipation and phi mass, max |div u| of each
form, and the largest tracer displacement between the two (tracers are
matched by tag).

    python -m mhit36_mixedp_amd.precision --n 256 --steps 1000 --phi 1 --ndrops 64 \\
        --radius 0.2 --ntracers 1048576 --dt 2e-4 --mu 0.002
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

from . import _abi
from .api import Solver


def _solver(prec, a):
    sigma = (2 * a.radius) / a.we if a.phi else 1.0
    s = Solver(nx=a.n, poisson_precision=prec, phiflag=int(a.phi), partflag=int(a.ntracers > 0),
               dt=a.dt, mu=a.mu, rho=1.0, sigma=sigma, epsr=1.0, k2_mode=a.k2)
    s.init_random_velocity(a.seed, kmax=8.0, urms=1.0)
    if a.phi:
        s.init_drops(a.ndrops, a.radius, a.seed)
    if a.ntracers:
        s.init_tracers_uniform(a.ntracers, a.seed)
    return s


def _rel(x, ref):
    return abs(x - ref) / abs(ref) if ref != 0 else abs(x)


def _periodic_dist(a, b, L=2 * np.pi):
    d = np.abs(a - b)
    return np.minimum(d, L - d)


def compare(a) -> dict:
    """Advance both forms a.steps steps; errors of the FP32 form at the end
    and the largest relative KE error seen at the checkpoints."""
    s32, s64 = _solver(_abi.POISSON_FP32, a), _solver(_abi.POISSON_FP64, a)
    checkpoints = sorted({max(1, a.steps // 10 * i) for i in range(1, 11)} | {a.steps})
    done, ke_max = 0, 0.0
    mass0 = s64.diag()["phi_mass"] if a.phi else 0.0
    for c in checkpoints:
        s32.step(c - done)
        s64.step(c - done)
        done = c
        ke_max = max(ke_max, _rel(s32.diag()["ke"], s64.diag()["ke"]))
    d32, d64 = s32.diag(), s64.diag()
    out = {
        "n": a.n, "steps": a.steps, "phase": bool(a.phi), "tracers": a.ntracers,
        "k2_mode": "modified" if a.k2 else "continuous",
        "ke_rel_err": _rel(d32["ke"], d64["ke"]),
        "ke_rel_err_max": ke_max,
        "dissipation_rel_err": _rel(d32["dissipation"], d64["dissipation"]),
        "maxdiv_fp32": d32["maxdiv"],
        "maxdiv_fp64": d64["maxdiv"],
        "nan_count": d32["nan_count"] + d64["nan_count"],
    }
    if a.phi:
        out["phi_mass_rel_err"] = _rel(d32["phi_mass"], d64["phi_mass"])
        out["phi_mass_drift_fp32"] = _rel(d32["phi_mass"], mass0)
        out["phi_mass_drift_fp64"] = _rel(d64["phi_mass"], mass0)
    if a.ntracers:
        t32, t64 = s32.download_tracers(), s64.download_tracers()
        o32, o64 = np.argsort(t32["tag"]), np.argsort(t64["tag"])
        dist = np.sqrt(sum(_periodic_dist(t32[k][o32], t64[k][o64]) ** 2 for k in ("x", "y", "z")))
        out["tracer_pos_err_max"] = float(dist.max())
        out["tracer_pos_err_rms"] = float(np.sqrt(np.mean(dist ** 2)))
    s32.close()
    s64.close()
    return out


def parser():
    ap = argparse.ArgumentParser(description="FP32 vs FP64 Poisson error on the GPU")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--phi", type=int, default=0)
    ap.add_argument("--ndrops", type=int, default=64)
    ap.add_argument("--radius", type=float, default=0.2)
    ap.add_argument("--we", type=float, default=1.0)
    ap.add_argument("--ntracers", type=int, default=0)
    ap.add_argument("--dt", type=float, default=2e-4)
    ap.add_argument("--mu", type=float, default=0.002)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--k2", type=int, default=_abi.K2_CONTINUOUS,
                    help="0: continuous k^2 (reference), 1: modified k'^2")
    return ap


def main(argv=None):
    a = parser().parse_args(argv)
    print(json.dumps(compare(a)), flush=True)


if __name__ == "__main__":
    sys.exit(main())
