"""GPU timestep parity against the CPU oracle (FP64 restatement of
multi/main.f90:456-1007) on identical seeded inputs.

Bars: FP64 Poisson 1e-12 relative after one step and 1e-10 after many
(the FP64 mode's tolerance against the reference's FP64 path); FP32
Poisson against the oracle's FP32 form 1e-6 after one step.  Tracer positions and phi use
the same bars."""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu


def make_pair(oracle_mod, n, prec, phiflag=0, partflag=0, ndrops=0, radius=0.5, ntr=0,
              seed=1, kmax=4.0, **kw):
    from mhit36_mixedp_amd import FIELD_PHI, Solver
    kw.setdefault("dt", 5e-4)
    kw.setdefault("mu", 0.006)
    s = Solver(nx=n, poisson_precision=prec, phiflag=phiflag, partflag=partflag, **kw)
    s.init_random_velocity(seed, kmax=kmax, urms=1.0)
    o = oracle_mod.Oracle(nx=n, poisson_precision=prec, phiflag=phiflag, partflag=partflag, **kw)
    u, v, w = s.velocity()
    o.set_field(1, u); o.set_field(2, v); o.set_field(3, w)
    if phiflag:
        if ndrops:
            s.init_drops(ndrops, radius, seed)
        else:
            s.init_drop()
        o.set_field(FIELD_PHI, s.download_field(FIELD_PHI))
    if ntr:
        s.init_tracers_uniform(ntr, seed)
        t = s.download_tracers()
        o.set_tracers(t["x"], t["y"], t["z"])
    return s, o


def compare(s, o, phi=False, tr=False):
    errs = {}
    for f, name in ((1, "u"), (2, "v"), (3, "w")):
        errs[name] = relerr(s.download_field(f), o.get_field(f))
    errs["p"] = relerr(s.download_field(4), o.get_field(4))
    if phi:
        errs["phi"] = relerr(s.download_field(5), o.get_field(5))
    if tr:
        t = s.download_tracers()
        # the GPU re-orders tracers by cell; tags 1..N identify them
        order = np.argsort(t["tag"])
        assert np.array_equal(t["tag"][order], np.arange(1, t["tag"].size + 1))
        t = {k: a[order] for k, a in t.items()}
        ox, oy, oz, ou, ov, ow = o.get_tracers()
        errs["xp"] = max(relerr(t["x"], ox), relerr(t["y"], oy), relerr(t["z"], oz))
        errs["up"] = max(relerr(t["up"], ou), relerr(t["vp"], ov), relerr(t["wp"], ow))
    return errs


def test_initial_field_divergence_free():
    from mhit36_mixedp_amd import Solver
    with Solver(nx=32) as s:
        s.init_random_velocity(3, kmax=6.0, urms=1.0)
        d = s.diag()
    print("[init] ", d, flush=True)
    assert d["maxdiv"] < 1e-12
    assert abs(np.sqrt(2 * d["ke"] / 3) - 1.0) < 1e-12
    assert abs(d["umean"]) == 0.0


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
def test_ns_one_step(prec, oracle_mod):
    print(f"[ns 1 step] prec={prec}", flush=True)
    s, o = make_pair(oracle_mod, 32, prec)
    s.step(1)
    assert o.step(1) == 0
    e = compare(s, o)
    print(e, flush=True)
    tol = 1e-12 if prec == 0 else 1e-6
    assert max(e.values()) < tol, e
    ds, do = s.diag(), o.diag()
    assert ds["gumax"] == pytest.approx(do["gumax"], rel=1e-12 if prec == 0 else 1e-6)
    assert ds["alpha"] == 1.5 and ds["beta"] == 0.5


def test_ns_many_steps_fp64(oracle_mod):
    print("[ns 40 steps fp64]", flush=True)
    s, o = make_pair(oracle_mod, 32, 0)
    s.step(40)
    assert o.step(40) == 0
    e = compare(s, o)
    print(e, flush=True)
    assert max(e.values()) < 1e-10, e
    ds, do = s.diag(), o.diag()
    for k in ("ke", "dissipation", "gumax"):
        assert ds[k] == pytest.approx(do[k], rel=1e-10), k


@pytest.mark.parametrize("path", ["march", "cell"])
@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
def test_multiphase_tracers_steps(prec, path, oracle_mod, monkeypatch):
    """Both GPU stencil paths: the fused z-marching passes (default) and the
    cell-per-thread kernels split as the reference's loops."""
    print(f"[multiphase 10 steps] prec={prec} path={path}", flush=True)
    monkeypatch.setenv("MHIT_STENCIL_PATH", path)
    s, o = make_pair(oracle_mod, 32, prec, phiflag=1, partflag=1, ndrops=3, radius=0.6,
                     ntr=2000, sigma=0.5)
    s.step(1)
    assert o.step(1) == 0
    e1 = compare(s, o, phi=True, tr=True)
    print("1 step", e1, flush=True)
    s.step(9)
    assert o.step(9) == 0
    e = compare(s, o, phi=True, tr=True)
    print("10 steps", e, flush=True)
    tol1 = 1e-12 if prec == 0 else 1e-6
    tol = 1e-10 if prec == 0 else 1e-5
    assert max(e1.values()) < tol1, e1
    assert max(e.values()) < tol, e
    ds, do = s.diag(), o.diag()
    assert ds["phi_mass"] == pytest.approx(do["phi_mass"], rel=1e-12)


def test_tgv_initial_condition():
    """Taylor-Green (main.f90:324-337): sampled over whole periods
    <u^2> = <v^2> = 1/8, so KE = 1/8; w = 0; discretely divergence free
    is not expected (centred staggering), but the mean is zero."""
    from mhit36_mixedp_amd import Solver
    with Solver(nx=32) as s:
        s.init_taylor_green()
        d = s.diag()
        u, v, w = s.velocity()
    assert d["ke"] == pytest.approx(0.125, rel=1e-12)
    assert np.all(w == 0)
    assert abs(u.mean()) < 1e-15 and abs(v.mean()) < 1e-15


def test_cfl_abort():
    """cou > 7 stops the run after the step that exceeded it (main.f90:1006)."""
    from mhit36_mixedp_amd import CFLStop, Solver
    with Solver(nx=16, dt=0.5) as s:
        s.init_random_velocity(5, kmax=3.0, urms=20.0)
        with pytest.raises(CFLStop):
            s.step(3)
        d = s.diag()
        assert d["step"] == 1
        assert d["cou"] > 7
        with pytest.raises(CFLStop):
            s.step(1)


@pytest.mark.parametrize("n", [96, 24])
def test_multiphase_partial_tiles_and_chunks(n, oracle_mod):
    """Sizes that are not multiples of the 64 x 8 / 64 x 16 tiles and give
    several 32-plane z chunks (96) or fewer planes than a chunk (24)."""
    print(f"[multiphase n={n}]", flush=True)
    s, o = make_pair(oracle_mod, n, 0, phiflag=1, partflag=1, ndrops=4, radius=0.5, ntr=500,
                     sigma=0.4, kmax=5.0)
    s.step(3)
    assert o.step(3) == 0
    e = compare(s, o, phi=True, tr=True)
    print(e, flush=True)
    assert max(e.values()) < 1e-11, e


@pytest.mark.parametrize("interp", [0, 1], ids=["reference", "trilinear"])
def test_tracers_on_cell_faces_and_domain_edges(oracle_mod, interp):
    """Tracers placed exactly on grid lines, at 0 and just below 2 pi on each
    axis (the wrap of the interpolation, particles.f90:78-122, and of the
    position, main.f90:501-508), stepped against the oracle."""
    print(f"[tracers] edges interp={interp}", flush=True)
    n = 24
    L = 2 * np.pi
    dx = L / n
    edge = [0.0, dx, 0.5 * dx, L - 1e-13, L - dx, L - 0.5 * dx, 3 * dx, L / 2]
    pts = np.array([(a, b, c) for a in edge for b in edge[:4] for c in edge[2:6]])
    s, o = make_pair(oracle_mod, n, 0, partflag=1, interp_mode=interp)
    s.upload_tracers(pts[:, 0], pts[:, 1], pts[:, 2])
    o.set_tracers(pts[:, 0], pts[:, 1], pts[:, 2])
    for _ in range(3):
        s.step(1)
        assert o.step(1) == 0
    e = compare(s, o, tr=True)
    t = s.download_tracers()
    assert np.all((t["x"] >= 0) & (t["x"] < L)) and np.all((t["z"] >= 0) & (t["z"] < L))
    assert max(e.values()) < 1e-12, e
    s.close()
    o.close()


def test_tracer_flag_with_no_tracers(oracle_mod):
    """partflag = 1 with an empty tracer set steps like partflag = 0."""
    print("[tracers] none", flush=True)
    s, o = make_pair(oracle_mod, 16, 0, partflag=1)
    s.upload_tracers(np.zeros(0), np.zeros(0), np.zeros(0))
    assert s.download_tracers()["x"].size == 0
    s.step(3)
    assert o.step(3) == 0
    e = compare(s, o)
    assert max(e.values()) < 1e-12, e
    s.close()
    o.close()


def test_scalars_match_diag():
    """mhit_scalars returns the step's scalars of mhit_diag_compute without
    the field pass (its diagnostic members stay zero)."""
    from mhit36_mixedp_amd import Solver
    print("[step] scalars vs diag, 24^3 multiphase", flush=True)
    with Solver(nx=24, poisson_precision=1, phiflag=1, partflag=0, dt=5e-4, mu=0.006, sigma=0.5) as s:
        s.init_random_velocity(5, kmax=4.0, urms=1.0)
        s.init_drops(2, 0.5, 5)
        s.step(3)
        a, d = s.scalars(), s.diag()
    for k in ("step", "alpha", "beta", "gumax", "cou", "umean", "vmean", "wmean"):
        assert a[k] == d[k], k
    assert a["step"] == 3 and a["alpha"] == 1.5 and a["beta"] == 0.5
    for k in ("ke", "dissipation", "maxdiv", "phi_mass", "nan_count"):
        assert a[k] == 0, k
