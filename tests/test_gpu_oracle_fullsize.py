"""GPU timestep against the CPU oracle (FP64 restatement of
multi/main.f90:456-1007) at production sizes, on identical seeded inputs:
the multiphase + tracer shape at 256^3 (64 drops of R = 0.2, 1,048,576
tracers) for 2 steps and at 512^3 (512 drops of R = 0.1, 4,194,304 tracers)
for 1 step.

Bars: FP64 Poisson 1e-12 relative to each field's largest value, FP32
Poisson 1e-6 against the oracle's FP32 form; velocity, pressure, phi,
tracer positions and velocities (matched by tag) and the diagnostic
scalars.  One FP32 exception: max |div u| after the projection is the
difference of two nearly equal fields, div u* and dt/rho lap_h p, and the
FP32 transforms' rounding of p (the two FFTs differ in algorithm, about
4e-7 relative in p) is amplified by lap_h toward the grid scale; its bar
in FP32 is 1e-5 (measured 4.2e-6 at 256^3).  The oracle's sums are
compensated, so the diagnostic sums are checked against near-exact values.
The oracle takes seconds per 512^3 step on the host's cores."""
import pytest

from test_gpu_step import compare, make_pair

pytestmark = pytest.mark.gpu

DIAG_KEYS = ("ke", "dissipation", "maxdiv", "phi_mass", "gumax", "umean", "vmean", "wmean")


def _run(oracle_mod, n, prec, steps, ndrops, radius, ntr, dt):
    print(f"[oracle parity] n={n} prec={prec} {steps} step(s), {ndrops} drops, {ntr} tracers",
          flush=True)
    s, o = make_pair(oracle_mod, n, prec, phiflag=1, partflag=1, ndrops=ndrops, radius=radius,
                     ntr=ntr, seed=2024, kmax=8.0, dt=dt, mu=0.002, rho=1.0,
                     sigma=2 * radius, epsr=1.0)
    try:
        s.step(steps)
        assert o.step(steps) == 0
        e = compare(s, o, phi=True, tr=True)
        ds, do = s.diag(), o.diag()
        for k in DIAG_KEYS:
            scale = abs(do[k]) if k not in ("umean", "vmean", "wmean") else 1.0
            e["diag_" + k] = abs(ds[k] - do[k]) / scale
    finally:
        s.close()
        o.close()
    print(e, flush=True)
    tol = 1e-12 if prec == 0 else 1e-6
    if prec == 1:
        assert e.pop("diag_maxdiv") < 1e-5, e
    assert max(e.values()) < tol, e


@pytest.mark.timeout(300)
@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
def test_oracle_parity_256_multiphase_tracers(oracle_mod, prec):
    _run(oracle_mod, 256, prec, 2, 64, 0.2, 1048576, 2e-4)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
def test_oracle_parity_512_multiphase_tracers(oracle_mod, prec):
    _run(oracle_mod, 512, prec, 1, 512, 0.1, 4194304, 1e-4)
