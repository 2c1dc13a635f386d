#!/usr/bin/env python3
"""MHIT36 single-GPU timestep benchmark on MI355X.

Workload (default): 512^3 multiphase box -- random-phase velocity
(|k| <= 8, u_rms = 1, seeded), ACDI phase field with 512 seeded drops of
radius 0.1 (eps = dx), CSF surface tension at We = rho u_rms^2 (2R)/sigma
= 1, 4,194,304 uniform tracers, AB2 with dt = 1e-4, nu = 0.002.  All inputs
are generated on the device before the timed region.

Prints ONE JSON line: ms per timestep of the FP32-Poisson form (the
"mixedp" form of the reference) with the FP64-Poisson form beside it, the
whole-step HBM rate over the algorithmic bytes, a roofline entry for the
dominant kernel (HIP events on the library's stream) and the CPU oracle
timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def step_bytes_fused(n: int, ntr: int, phase: bool, fp32: bool) -> float:
    """HBM bytes per step of this pass structure with every array read and
    written once per pass, plus 144 B per tracer.  The reference's pass
    structure needs 208/248 B per cell (single phase, FP32/FP64 Poisson) and
    272/312 with the phase field; here the NS pass also writes div u* (4/8 B)
    and the x R2C pass reads it instead of u*, v*, w* (24 B): 192/240 and
    256/304 B per cell."""
    per_cell = (256.0 if fp32 else 304.0) if phase else (192.0 if fp32 else 240.0)
    return per_cell * n ** 3 + 144.0 * ntr


def make_solver(args, prec):
    from mhit36_mixedp_amd import Solver
    sigma = 1.0 * 1.0 ** 2 * (2 * args.radius) / args.we
    s = Solver(nx=args.n, poisson_precision=prec, phiflag=1, partflag=1, dt=args.dt, mu=args.mu,
               rho=1.0, sigma=sigma, epsr=1.0, npart=args.ntracers)
    s.init_random_velocity(args.seed, kmax=8.0, urms=1.0)
    s.init_drops(args.ndrops, args.radius, args.seed)
    s.init_tracers_uniform(args.ntracers, args.seed)
    return s


def time_steps(s, warmup, steps):
    s.step(warmup)
    s.sync()
    t0 = time.perf_counter()
    s.step(steps)           # returns after a stream synchronize
    t1 = time.perf_counter()
    return (t1 - t0) * 1e3 / steps


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_512_multiphase.json")


def measured_traffic(stage: str, n: int):
    """HBM bytes per launch of a stage's kernel from the committed rocprofv3
    --pmc summary (FETCH_SIZE and WRITE_SIZE, each scaled by what the counter
    reads for a plain copy of known size), or None if that file does not
    cover this stage at this size."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("n") != n:
        return None
    k = t.get("stages", {}).get(stage)
    return None if k is None else float(k["read_bytes"] + k["write_bytes"])


def host_cpu():
    """The host processor model and the logical CPUs this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return {"model": model, "nproc": usable, "cpu_count": os.cpu_count()}


def cpu_baseline(args, fields, tracers):
    """Time the CPU oracle on one step of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = min(16, os.cpu_count() or 1)
    oracle.set_threads(threads)
    sigma = (2 * args.radius) / args.we
    o = oracle.Oracle(nx=args.n, poisson_precision=1, phiflag=1, partflag=1, dt=args.dt, mu=args.mu,
                      rho=1.0, sigma=sigma, epsr=1.0)
    for f, a in fields.items():
        o.set_field(f, a)
    o.set_tracers(*tracers)
    t0 = time.perf_counter()
    rc = o.step(args.cpu_steps)
    t1 = time.perf_counter()
    o.close()
    if rc != 0:
        raise RuntimeError(f"oracle step failed: {rc}")
    return (t1 - t0) * 1e3 / args.cpu_steps, threads


def box_comparison(args, n, prec, mu, dt, cpu_runs):
    """Single-phase NS box of size n from a random-phase |k| <= 8 field:
    GPU ms/step over 100 steps; CPU oracle median ms/step over the timed
    steps after 5 warm-up steps, for each (threads, timed steps) in
    cpu_runs."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle
    from mhit36_mixedp_amd import Solver
    kw = dict(nx=n, poisson_precision=prec, phiflag=0, partflag=0, dt=dt, mu=mu)
    with Solver(**kw) as s:
        s.init_random_velocity(args.seed, kmax=8.0, urms=1.0)
        vel = s.velocity()
        s.step(5)
        s.sync()
        t0 = time.perf_counter()
        s.step(100)
        gpu_ms = (time.perf_counter() - t0) * 1e3 / 100
    res = {"workload": f"{n}^3 single-phase NS, {'FP32' if prec else 'FP64'} Poisson, nu={mu}, dt={dt}",
           "gpu_ms_per_step": round(gpu_ms, 4)}
    for th, steps in cpu_runs:
        oracle.set_threads(th)
        o = oracle.Oracle(**kw)
        for f, a in zip((1, 2, 3), vel):
            o.set_field(f, a)
        o.step(5)
        ts = []
        for _ in range(steps):
            t0 = time.perf_counter()
            o.step(1)
            ts.append((time.perf_counter() - t0) * 1e3)
        o.close()
        res[f"cpu_ms_per_step_{th}_threads"] = round(float(np.median(ts)), 3)
    return res


def small_box_comparison(args):
    """The reference's CPU case (64^3 single-phase NS, FP64 throughout,
    nu = 0.006, dt = 5e-4) on 1 thread and on all threads used for the
    baseline, 10 timed CPU steps each; and the 256^3 single-phase case
    (FP32 Poisson, nu = 0.002, dt = 2e-4) for a rate comparison, 3 timed
    CPU steps on all threads."""
    threads_all = min(16, os.cpu_count() or 1)
    res = box_comparison(args, 64, 0, 0.006, 5e-4, [(1, 10), (threads_all, 10)])
    res["c2_256"] = box_comparison(args, 256, 1, 0.002, 2e-4, [(threads_all, 3)])
    return res


ENSEMBLE_SIZES = (64, 96, 128)


def one_size_rates(args):
    """Box-steps/s of 16 boxes of one size (every other box multiphase),
    batched, for each size of the ensemble workload: the case where one
    size class alone has to fill the GPU."""
    from mhit36_mixedp_amd import Ensemble, default_params
    out = {}
    for n in ENSEMBLE_SIZES:
        params = [default_params(nx=n, phiflag=i % 2, partflag=0, poisson_precision=1, dt=5e-4,
                                 mu=0.004 + 0.0005 * (i % 9), rho=1.0, sigma=2 * 0.4, epsr=1.0)
                  for i in range(16)]
        with Ensemble(params) as ens:
            for i, b in enumerate(ens.boxes):
                b.init_random_velocity(args.seed + i, kmax=4.0, urms=1.0)
                if i % 2:
                    b.init_drops(1 + i % 6, 0.4, args.seed + i)
            ens.step(args.warmup)
            t0 = time.perf_counter()
            ens.step(args.ensemble_steps)
            dt = (time.perf_counter() - t0) / args.ensemble_steps
        out[f"16x{n}^3"] = round(len(params) / dt, 1)
    return out


def ensemble_leg(args, batched_only=False):
    """Ensemble workload: 16 boxes each of 64^3, 96^3 and 128^3 (51,904,512
    cells), every other box multiphase, each with its own seed, viscosity
    and drop count.  Returns box-steps/s stepped together as one captured
    graph per step -- batched (boxes of one size share every launch) and
    with one graph branch per box -- and stepped one box at a time, with
    the algorithmic HBM rate of the batched form."""
    from mhit36_mixedp_amd import Ensemble, default_params
    params, phase = [], []
    for i in range(16 * len(ENSEMBLE_SIZES)):
        n = ENSEMBLE_SIZES[i // 16]
        ph = i % 2
        params.append(default_params(nx=n, phiflag=ph, partflag=0, poisson_precision=1, dt=5e-4,
                                     mu=0.004 + 0.0005 * (i % 9), rho=1.0, sigma=2 * 0.4,
                                     epsr=1.0))
        phase.append(ph)
    cells = sum(int(p.nx) ** 3 for p in params)
    nbytes = sum(step_bytes_fused(int(p.nx), 0, bool(ph), True) for p, ph in zip(params, phase))
    with Ensemble(params) as ens:
        for i, (b, ph) in enumerate(zip(ens.boxes, phase)):
            b.init_random_velocity(args.seed + i, kmax=4.0, urms=1.0)
            if ph:
                b.init_drops(1 + i % 6, 0.4, args.seed + i)
        ens.step(args.warmup)
        t0 = time.perf_counter()
        ens.step(args.ensemble_steps)
        t1 = time.perf_counter()
        together = (t1 - t0) / args.ensemble_steps          # s per ensemble step
        # the batched shapes tuned per grid size on the live state (restored),
        # then timed like the default shapes
        ens.autotune(steps=3)
        tuned_shapes = {f"{n}^3": ens.boxes[16 * k].tuning() for k, n in enumerate(ENSEMBLE_SIZES)}
        for t in tuned_shapes.values():
            t.pop("ms_per_step", None)
            t.pop("sort_every", None)
        ens.step(args.warmup)
        t0 = time.perf_counter()
        ens.step(args.ensemble_steps)
        t1 = time.perf_counter()
        tuned = (t1 - t0) / args.ensemble_steps
        if batched_only:
            nbox = len(params)
            return {"workload": "16x64^3 + 16x96^3 + 16x128^3 boxes, half multiphase, FP32 Poisson",
                    "boxes": nbox, "cells": cells, "steps": args.ensemble_steps,
                    "box_steps_per_s": round(nbox / together, 1),
                    "ms_per_ensemble_step": round(together * 1e3, 4),
                    "hbm_GBps": round(nbytes / together / 1e9, 1),
                    "autotuned_box_steps_per_s": round(nbox / tuned, 1),
                    "autotuned_shapes": tuned_shapes}
        ens.set_batched(False)
        ens.step(args.warmup)
        t0 = time.perf_counter()
        ens.step(args.ensemble_steps)
        t1 = time.perf_counter()
        branches = (t1 - t0) / args.ensemble_steps
        t0 = time.perf_counter()
        for b in ens.boxes:
            b.step(args.ensemble_steps)
        t1 = time.perf_counter()
        alone = (t1 - t0) / args.ensemble_steps
        nan = sum(b.diag()["nan_count"] for b in ens.boxes)
    if nan:
        raise RuntimeError("non-finite values in the ensemble")
    nbox = len(params)
    return {
        "workload": "16x64^3 + 16x96^3 + 16x128^3 boxes, half multiphase, FP32 Poisson",
        "boxes": nbox, "cells": cells, "steps": args.ensemble_steps,
        "box_steps_per_s": round(nbox / together, 1),
        "ms_per_ensemble_step": round(together * 1e3, 4),
        "hbm_GBps": round(nbytes / together / 1e9, 1),
        "autotuned_box_steps_per_s": round(nbox / tuned, 1),
        "autotuned_shapes": tuned_shapes,
        "branch_per_box_box_steps_per_s": round(nbox / branches, 1),
        "one_at_a_time_box_steps_per_s": round(nbox / alone, 1),
        "one_size_box_steps_per_s": one_size_rates(args),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--ndrops", type=int, default=512)
    ap.add_argument("--radius", type=float, default=0.1)
    ap.add_argument("--ntracers", type=int, default=4194304)
    ap.add_argument("--dt", type=float, default=1e-4)
    ap.add_argument("--mu", type=float, default=0.002)
    ap.add_argument("--we", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-Poisson form")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-copy-probe", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=20)
    ap.add_argument("--no-ensemble", action="store_true", help="skip the ensemble workload")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="run only the batched ensemble and print its JSON (for kernel traces)")
    ap.add_argument("--one-size-only", action="store_true",
                    help="run only the one-size ensembles (16 boxes each of 64^3, 96^3, 128^3) "
                         "and print their box-steps/s (for kernel traces)")
    ap.add_argument("--autotune", action="store_true",
                    help="let mhit_autotune pick the launch shapes before timing (default: the "
                         "library's size-based defaults, which the kernel trace under profiles/ uses)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: the time-stepper runs one box on one GPU; --gpus {args.gpus} is not supported "
              "(independent boxes shard trivially across GPUs; see the sharding notes).",
              file=sys.stderr)
        sys.exit(2)

    from mhit36_mixedp_amd import FIELD_PHI, FIELD_U, FIELD_V, FIELD_W, version
    if args.ensemble_only:
        log(f"[bench] {version()}  batched ensemble only")
        print(json.dumps({"ensemble": ensemble_leg(args, batched_only=True)}), flush=True)
        return
    if args.one_size_only:
        log(f"[bench] {version()}  one-size ensembles only")
        print(json.dumps({"one_size_box_steps_per_s": one_size_rates(args)}), flush=True)
        return
    log(f"[bench] {version()}  n={args.n} drops={args.ndrops} tracers={args.ntracers}")

    # --- FP32-Poisson form (the measured value) --------------------------------
    s = make_solver(args, prec=1)
    fields = tracers = None
    if not args.no_cpu_baseline:
        fields = {f: s.download_field(f) for f in (FIELD_U, FIELD_V, FIELD_W, FIELD_PHI)}
        t = s.download_tracers()
        tracers = (t["x"], t["y"], t["z"])
    copy_gbps = s.copy_probe(1 << 30, 10) if not args.no_copy_probe else None
    # the NS pass's mix: 7 FP64 read streams (u, v, w, phi, 3 rhs_o) and
    # 6 FP64 write streams (u*, v*, w*, 3 rhs) -- plus div u* in FP32
    stream_gbps = s.stream_probe(512 << 20, 7, 6, 5) if not args.no_copy_probe else None
    if copy_gbps:
        log(f"[bench] plain device copy: {copy_gbps:.0f} GB/s (read + write); "
            f"7-read/6-write streaming: {stream_gbps:.0f} GB/s")
    s_tuning = s.tuning()                     # the shapes timed below
    if args.autotune:
        s_tuning = s.autotune(steps=3)        # set-up, before the timed region
        log(f"[bench] autotune: {s_tuning}")
    ms32 = time_steps(s, args.warmup, args.steps)
    log(f"[bench] FP32 Poisson: {ms32:.3f} ms/step")
    s.profile_enable(True)
    s.step(args.profile_steps)
    prof32 = s.profile()
    s.profile_enable(False)
    d = s.diag()
    log(f"[bench] diag after {d['step']} steps: ke={d['ke']:.6e} gumax={d['gumax']:.4f} "
        f"cou={d['cou']:.3f} maxdiv={d['maxdiv']:.3e} phi_mass={d['phi_mass']:.6e} nan={d['nan_count']}")
    if d["nan_count"]:
        raise RuntimeError("non-finite values in the state")
    s.close()
    del s

    # --- FP64-Poisson form ------------------------------------------------------
    ms64 = None
    prof64 = None
    if not args.no_fp64:
        s = make_solver(args, prec=0)
        if args.autotune:
            log(f"[bench] autotune (FP64 Poisson): {s.autotune(steps=3)}")
        ms64 = time_steps(s, args.warmup, args.steps)
        log(f"[bench] FP64 Poisson: {ms64:.3f} ms/step")
        s.profile_enable(True)
        s.step(args.profile_steps)
        prof64 = s.profile()
        s.close()
        del s

    # --- ensemble of small boxes ----------------------------------------------------
    ens = None
    if not args.no_ensemble:
        ens = ensemble_leg(args)
        log(f"[bench] ensemble: {ens['box_steps_per_s']} box-steps/s together "
            f"({ens['autotuned_box_steps_per_s']} autotuned), "
            f"{ens['one_at_a_time_box_steps_per_s']} one box at a time, {ens['hbm_GBps']} GB/s")

    # --- per-stage table and roofline of the dominant kernel ----------------------
    log(f"[bench] per-stage (FP32 Poisson, {args.profile_steps} profiled steps):")
    for r in prof32:
        avg = r["ms"] / r["calls"]
        gbps = r["bytes_per_call"] / (avg * 1e-3) / 1e9 if avg > 0 else 0.0
        log(f"   {r['stage']:<14} {avg:8.3f} ms  {r['bytes_per_call'] / 1e9:7.3f} GB  {gbps:8.1f} GB/s")
    if prof64:
        log(f"[bench] per-stage (FP64 Poisson, {args.profile_steps} profiled steps):")
        for r in prof64:
            avg = r["ms"] / r["calls"]
            gbps = r["bytes_per_call"] / (avg * 1e-3) / 1e9 if avg > 0 else 0.0
            log(f"   {r['stage']:<14} {avg:8.3f} ms  {r['bytes_per_call'] / 1e9:7.3f} GB  {gbps:8.1f} GB/s")
    dom = max(prof32, key=lambda r: r["ms"])
    dom_avg_ms = dom["ms"] / dom["calls"]
    dom_gbps = dom["bytes_per_call"] / (dom_avg_ms * 1e-3) / 1e9

    step_bytes32 = step_bytes_fused(args.n, args.ntracers, True, True)
    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] CPU oracle on the same workload ...")
        cpu_ms, threads = cpu_baseline(args, fields, tracers)
        log(f"[bench] CPU oracle: {cpu_ms:.1f} ms/step on {threads} threads")
        cpu = {"value": round(cpu_ms, 3), "unit": "ms/timestep", "cores": threads, "kind": "port",
               "sample": f"{args.cpu_steps} timestep(s) of the same {args.n}^3 multiphase workload "
                         f"({args.ndrops} drops, {args.ntracers} tracers, FP32 Poisson) on the CPU "
                         "FP64 restatement (oracle/), OpenMP",
               "host": host_cpu()}

    small = None
    if not args.no_cpu_baseline:
        small = small_box_comparison(args)
        log(f"[bench] 64^3 NS: {small}")

    out = {
        "metric": "ms/timestep & HBM GB/s at 512^3 multiphase, FP32 vs FP64 Poisson, 1 MI355X",
        "value": round(ms32, 4),
        "unit": "ms/timestep",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms32, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 state, f32 Poisson",
        "data": "synthetic (seeded on device)",
        "config": {
            "workload": f"{args.n}^3 multiphase: {args.ndrops} drops R={args.radius}, "
                        f"{args.ntracers} tracers, dt={args.dt}, nu={args.mu}, We={args.we}",
            "n": args.n, "ndrops": args.ndrops, "radius": args.radius, "ntracers": args.ntracers,
            "poisson": "fp32", "parallelism": "single GPU",
        },
        "hbm_GBps": round(step_bytes32 / (ms32 * 1e-3) / 1e9, 1),
        "step_bytes": step_bytes32,
        "fp64_poisson_ms_per_step": round(ms64, 4) if ms64 else None,
        "fp64_poisson_hbm_GBps": round(step_bytes_fused(args.n, args.ntracers, True, False)
                                       / (ms64 * 1e-3) / 1e9, 1) if ms64 else None,
        "fp32_over_fp64_rate": round(ms64 / ms32, 4) if ms64 else None,
        "roofline": {
            "kernel": dom["stage"],
            "bound": "hbm",
            "achieved": round(dom_gbps, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(dom_gbps / HBM_PEAK_GBPS, 4),
            "traffic": measured_traffic(dom["stage"], args.n),
            "avg_ms": round(dom_avg_ms, 4),
            "bytes_per_launch": dom["bytes_per_call"],
            "measured_copy_GBps": round(copy_gbps, 1) if copy_gbps else None,
            "measured_stream_mix_GBps": round(stream_gbps, 1) if stream_gbps else None,
            "stream_mix": "pointwise kernel with the NS pass's 7 FP64 read + 6 FP64 write streams",
        },
        "stages": {r["stage"]: round(r["ms"] / r["calls"], 4) for r in prof32},
        "fp64_poisson_stages": ({r["stage"]: round(r["ms"] / r["calls"], 4) for r in prof64}
                                if prof64 else None),
        "tuning": s_tuning,
        "ensemble": ens,
        "small_box": small,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
