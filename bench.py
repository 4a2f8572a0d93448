#!/usr/bin/env python3
"""Benchmark of the particle time step (move + collide + sample) on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload expansion]

A step is one uniGasCloud::evolve() equivalent over the whole case.  The
default workload is the 3-D nozzle expansion of argon from a 1 bar, 300 K
reservoir through a 2 mm throat towards a 1 Pa background: 200x100x100
planar-faced hexes (2e6 cells), 1e8 simulators (50 per cell, cell-weighted),
hybrid USP/DSMC switching on the local Knudsen number, dt = 1e-8 s, started
from the isentropic expansion, FP32 particle state with FP64 sampler /
(sigma_T c_r)max / NTC accumulators.  value = particle-steps/s = sum over
timed steps of the simulator count / wall time, with the inputs resident in
HBM, timed after 300 steps (three hybrid decompositions and more).  Prints
one JSON line with the per-phase HIP-event timings, the roofline of the
dominant phase, an all-FP64 line, the mixed-regime nozzle (uniform fill at
1e-3 of the reservoir density), the batched/served sweep, the FP32-vs-FP64
field comparison and a CPU baseline (the FP64 OpenMP oracle on a bounded
sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s (move+collide+sample); batched vs plain; p50/p99 case latency"


COLLISION_DESC = {"hybrid": "hybrid USP/DSMC (NTC+VHS)", "dsmc": "DSMC (NTC+VHS)", "bgk": "USP everywhere",
                  "none": "collisionless"}


# HBM bytes per launch of each phase's kernels at scale 1, from rocprofv3
# --pmc FETCH_SIZE and WRITE_SIZE runs of this bench line (FETCH_SIZE scaled by
# 2 and WRITE_SIZE by 1, the factors a 2 GiB copy kernel reads; averages of the
# last 20 launches; the profile summary named beside each holds the tables).
PMC_TRAFFIC = {
    "expansion": {"profile": "profiles/nozzle_expansion_1e8_fp32.md",
                  "collide": (0.71 + 0.95 + 2.41 + 2.69) * 1e9,      # k_collide + k_bgk_conserve5
                  "cells": (9.59 + 4.56) * 1e9, "move": (4.24 + 2.07) * 1e9},
    "nozzle": {"profile": "profiles/nozzle_1e8_fp32.md",
               "collide": (0.78 + 0.63 + 2.00 + 1.56) * 1e9,      # k_collide + k_bgk_conserve5
               "cells": (12.01 + 4.71) * 1e9, "move": (4.31 + 1.61) * 1e9},
}


def build_workload(name, scale, collision="hybrid"):
    """(spec, init, desc): init(case) gives a new case its starting state."""
    from unigasfoam_amd import workloads
    if name == "expansion":
        nx, ny, nz = int(200 * scale), int(100 * scale), int(100 * scale)
        spec, state = workloads.nozzle_expansion(nx=nx, ny=ny, nz=nz, per_cell=50)
        desc = (f"3-D argon nozzle expansion from a 1 bar, 300 K reservoir through a 2 mm throat towards a 1 Pa "
                f"background, {nx}x{ny}x{nz} hex cells, {nx * ny * nz * 50:.3g} simulators (50/cell, cell-weighted), "
                f"hybrid USP/DSMC (NTC+VHS), started from the isentropic expansion")
        return spec, (lambda c: workloads.apply_state(c, state)), desc
    if name == "nozzle":
        nx, ny, nz = int(200 * scale), int(100 * scale), int(100 * scale)
        spec, fill = workloads.nozzle(nx=nx, ny=ny, nz=nz, per_cell=50, collision_model=collision)
        desc = (f"3-D argon nozzle expansion, {nx}x{ny}x{nz} hex cells, {nx * ny * nz * 50:.3g} simulators "
                f"(50/cell), {COLLISION_DESC[collision]}")
    elif name == "cylinder":
        spec, fill = workloads.cylinder()
        desc = "Mach-10 N2/O2 over a cylinder, 400x200x1 hex, ~5e7 simulators, LB-VHS, inflow+deletion"
    elif name == "couette":
        spec, fill = workloads.couette()
        desc = "1-D N2 Couette, 500 cells, 1e5 simulators, LB-VHS, diffuse walls +/-250 m/s"
    elif name == "heat_bath":
        spec, fill = workloads.heat_bath()
        desc = "0-D argon heat bath, 10^3 cells, 2e4 simulators, VHS, specular box"
    else:
        raise SystemExit(f"unknown workload {name}")
    return spec, (lambda c: c.fill(**fill)), desc


def cpu_baseline(name, seconds, scale):
    """FP64 OpenMP oracle on a bounded sample of the same workload."""
    from oracle import orc
    from unigasfoam_amd import workloads
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    threads = orc.set_threads(threads)
    orc.set_fast_sums(True)
    if name == "nozzle":
        nx, ny, nz = 40, 20, 20
        spec, fill = workloads.nozzle(nx=nx, ny=ny, nz=nz, per_cell=50)
        init = (lambda c: c.fill(**fill))
        sample = f"nozzle {nx}x{ny}x{nz} cells, {nx * ny * nz * 50} simulators, same per-cell load and models"
    elif name == "expansion":
        nx, ny, nz = 40, 20, 20
        spec, state = workloads.nozzle_expansion(nx=nx, ny=ny, nz=nz, per_cell=50)
        init = (lambda c: workloads.apply_state(c, state))
        sample = (f"1 bar expansion {nx}x{ny}x{nz} cells, {nx * ny * nz * 50} simulators, same per-cell load, "
                  f"weighting and models")
    else:
        spec, init, _ = build_workload(name, scale)
        sample = "same workload"
    o = orc.OracleCase(spec, capacity=4 * (spec.mesh.n_cells * 60 + 4096))
    init(o)
    o.step(1)
    c0 = o.counters()["particle_steps"]
    t0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - t0 < seconds or steps < 3:
        o.step(1)
        steps += 1
    dt = time.perf_counter() - t0
    ps = o.counters()["particle_steps"] - c0
    orc.set_fast_sums(False)
    return {"value": ps / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"{sample}; {steps} steps in {dt:.1f} s (restated CPU baseline, not uniGasFoam)"}


# a served case has converged when the density error estimate of every cell
# holding gas is at most this (uniGasVolFields.C:1242-1252)
SWEEP_ERROR_TARGET = 0.025


def sweep_requests(n_cases, rate, seed=5):
    """Config C5: parametric Couette/Fourier cases with Poisson arrivals."""
    import numpy as np
    from unigasfoam_amd import workloads
    from unigasfoam_amd.serve import make_request
    params = workloads.sweep_params(n_cases, seed=seed)
    gaps = np.random.default_rng(seed + 1).exponential(1.0 / rate, n_cases)
    arrivals = np.cumsum(gaps) - gaps[0]
    reqs = []
    for i, p in enumerate(params):
        spec, fill = workloads.sweep_case(p)
        reqs.append(make_request(i, spec, fill, p["per_cell"], warmup=50, samples_per_cell=2000,
                                 error_target=SWEEP_ERROR_TARGET, check_every=10,
                                 arrival=float(arrivals[i])))
    return reqs


def served_summary(st, rate, plain_rate):
    c = st.summary()
    return {"arrival_rate_per_s": rate, "particle_steps_per_s": c["particle_steps_per_s"],
            "wall_s": round(c["wall_s"], 2),
            "converged_when": f"density error <= {SWEEP_ERROR_TARGET} in every cell",
            "p50_time_to_converged_sample_s": c["p50_latency_s"],
            "p99_time_to_converged_sample_s": c["p99_latency_s"],
            "max_concurrent_cases": c["max_concurrent_cases"],
            "vs_plain": c["particle_steps_per_s"] / plain_rate}


def sweep_leg(n_cases, rate, precision, sat_rate):
    from unigasfoam_amd.serve import Server, run_alone
    reqs = sweep_requests(n_cases, rate)
    print(f"[bench] sweep: {n_cases} cases, plain one at a time ...", file=sys.stderr, flush=True)
    alone = run_alone(reqs, precision=precision, use_graphs=True)
    print("[bench] sweep: batched (all cases admitted at once) ...", file=sys.stderr, flush=True)
    srv = Server(precision=precision, chunk=10, use_graphs=True, realtime=False)
    batched = srv.run(reqs)
    srv.close()
    print(f"[bench] sweep: served (Poisson arrivals at {rate}/s) ...", file=sys.stderr, flush=True)
    srv = Server(precision=precision, chunk=10, use_graphs=True, realtime=True)
    served = srv.run(reqs)
    srv.close()
    # the same cases arriving faster than they converge: the queue stays
    # non-empty, and the latencies include the wait for admission
    print(f"[bench] sweep: served (Poisson arrivals at {sat_rate}/s) ...", file=sys.stderr, flush=True)
    srv = Server(precision=precision, chunk=10, use_graphs=True, realtime=True)
    served_sat = srv.run(sweep_requests(n_cases, sat_rate))
    srv.close()
    a, b = alone.summary(), batched.summary()
    return {
        "workload": f"{n_cases} parametric 1-D Couette/Fourier cases (Kn 0.01-10, 50-2000 cells, 20-200 simulators/cell, "
                    f"Ar or N2 LB), 50 warm-up steps then sampled until the density error is at most {SWEEP_ERROR_TARGET} in every cell",
        "particle_steps": a["particle_steps"],
        "plain_one_at_a_time": {"particle_steps_per_s": a["particle_steps_per_s"], "wall_s": round(a["wall_s"], 2)},
        "batched": {"particle_steps_per_s": b["particle_steps_per_s"], "wall_s": round(b["wall_s"], 2),
                    "max_concurrent_cases": b["max_concurrent_cases"],
                    "vs_plain": b["particle_steps_per_s"] / a["particle_steps_per_s"]},
        "served": served_summary(served, rate, a["particle_steps_per_s"]),
        "served_saturated": served_summary(served_sat, sat_rate, a["particle_steps_per_s"]),
    }


def precision_leg():
    """FP32 state against all-FP64 on the GPU as paired ensembles of
    independent replicas (precision.compare_replicas): 1-D Couette flows of
    N2 with DSMC (LB) and with USP, 48 replicas of 1e5 simulators each."""
    from unigasfoam_amd import workloads
    from unigasfoam_amd.precision import compare_replicas
    out = {}
    for name, kw in (("couette_dsmc_lb", {}),
                     ("couette_usp", dict(collision_model="bgk", bgk_model="unifiedStochasticParticleSBGK"))):
        spec, fill = workloads.couette(n_cells=100, per_cell=1000, U_wall=250.0, **kw)
        spec.sampler_full = True
        out[name] = compare_replicas(spec, fill, warmup=1000, sample=2000, replicas=48)
    return {"workload": "1-D N2 Couette, walls +/-250 m/s, 100 cells x 1000 simulators, 48 paired replicas per "
                        "precision, 1000 steps then 2000 sampled steps; standard errors from the replicas",
            "fields": out}


def nozzle_line(args, steps=20):
    """The mixed-regime nozzle (a uniform fill at 1e-3 of the reservoir
    density, no cell weighting: DSMC and USP cells side by side), timed the
    same way as the headline."""
    from unigasfoam_amd.core import Case
    print("[bench] mixed-regime nozzle line ...", file=sys.stderr, flush=True)
    spec, init, desc = build_workload("nozzle", args.scale, args.collision)
    case = Case(spec, precision=args.precision)
    init(case)
    case.step(args.develop + args.warmup)
    case.sync()
    c0 = case.counters()
    case.timing(True)
    case.sync()
    t0 = time.perf_counter()
    case.step(steps)
    case.sync()
    el = time.perf_counter() - t0
    c1 = case.counters()
    phases, tsteps = case.phase_times()
    case.close()
    ps = c1["particle_steps"] - c0["particle_steps"]
    return {"workload": desc, "value": ps / el, "unit": "particle-steps/s", "ms_per_step": el / steps * 1e3,
            "steps": steps, "untimed_steps": args.develop + args.warmup,
            "phases_ms": {p: round(ms / max(tsteps, 1), 4) for p, ms in phases.items()},
            "counters_per_step": {k: (c1[k] - c0[k]) / steps for k in ("candidates", "collisions", "bgk_collisions")},
            "dsmc_cells": c1["n_dsmc_cells"]}


def fp64_line(args, spec, init, steps=20):
    """The same workload and developed state with all-FP64 particle state
    (the precision whose results equal the oracle bit for bit)."""
    from unigasfoam_amd.core import Case
    print("[bench] FP64 line ...", file=sys.stderr, flush=True)
    case = Case(spec, precision="fp64")
    init(case)
    case.step(args.develop + args.warmup)
    case.sync()
    c0 = case.counters()
    t0 = time.perf_counter()
    case.step(steps)
    case.sync()
    el = time.perf_counter() - t0
    c1 = case.counters()
    case.close()
    ps = c1["particle_steps"] - c0["particle_steps"]
    return {"value": ps / el, "unit": "particle-steps/s", "ms_per_step": el / steps * 1e3, "steps": steps,
            "dtype": "f64", "untimed_steps": args.develop + args.warmup}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--develop", type=int, default=300,
                    help="steps run from the fill before the warm-up, so the timed steps see a developed flow "
                         "(DSMC cells, collisions) rather than the initial state")
    ap.add_argument("--no-fp64-line", action="store_true", help="skip the all-FP64 timing of the same workload")
    ap.add_argument("--workload", default="expansion",
                    help="expansion (the 1 bar nozzle expansion, default), nozzle (mixed-regime nozzle), cylinder, "
                         "couette, heat_bath")
    ap.add_argument("--no-nozzle-line", action="store_true", help="skip the mixed-regime nozzle line")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--scale", type=float, default=1.0, help="linear scale of the nozzle mesh")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--probe", action="store_true", help="also run the stream-copy HBM probe")
    ap.add_argument("--collision", default="hybrid", choices=["hybrid", "dsmc", "bgk", "none"],
                    help="collision model of the nozzle workload")
    ap.add_argument("--sweep-cases", type=int, default=256, help="cases of the batched/served sweep (0: skip)")
    ap.add_argument("--sweep-rate", type=float, default=20.0, help="arrival rate of the served sweep (cases/s)")
    ap.add_argument("--sweep-saturated-rate", type=float, default=2000.0,
                    help="arrival rate of the second served leg, above what the batch can retire (cases/s)")
    ap.add_argument("--no-precision", action="store_true", help="skip the FP32-vs-FP64 field comparison")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.stderr.write("bench.py: only --gpus 1 is supported (single-GPU build)\n")
        return 2

    from unigasfoam_amd import roofline
    from unigasfoam_amd.core import Case, probe_copy

    spec, init, desc = build_workload(args.workload, args.scale, args.collision)
    print(f"[bench] workload: {desc}", file=sys.stderr, flush=True)
    t = time.perf_counter()
    case = Case(spec, precision=args.precision)
    init(case)
    case.sync()
    print(f"[bench] setup {time.perf_counter() - t:.1f} s, {case.count()} simulators", file=sys.stderr, flush=True)
    if args.develop > 0:
        t = time.perf_counter()
        case.step(args.develop)
        case.sync()
        print(f"[bench] developed {args.develop} steps in {time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
    case.step(args.warmup)
    case.sync()
    print("[bench] warmup done", file=sys.stderr, flush=True)

    c0 = case.counters()
    case.timing(True)
    case.sync()
    t0 = time.perf_counter()
    case.step(args.steps)
    case.sync()
    t1 = time.perf_counter()
    c1 = case.counters()
    phases, tsteps = case.phase_times()
    case.timing(False)
    elapsed = t1 - t0
    psteps = c1["particle_steps"] - c0["particle_steps"]
    value = psteps / elapsed
    R = 4 if args.precision == "fp32" else 8
    n_avg = psteps / args.steps
    nc = spec.mesh.n_cells
    cand = (c1["candidates"] - c0["candidates"]) / args.steps
    coll = (c1["collisions"] - c0["collisions"]) / args.steps
    n_bgk = (c1["bgk_collisions"] - c0["bgk_collisions"]) / args.steps
    bgk_on = spec.collision_model in ("bgk", "hybrid") and spec.bgk_model != "noBGKCollision"
    kw = dict(rows=roofline.sampled_rows(spec.sampler_full, len(spec.species)), bgk=bgk_on,
              hybrid=spec.collision_model == "hybrid")

    def pbytes(p):
        return roofline.phase_bytes(p, R, n_avg, nc, cand, coll, n_bgk, **kw)

    per_phase = {}
    for p, ms in phases.items():
        avg = ms / max(tsteps, 1)
        b = pbytes(p)
        per_phase[p] = {"ms": round(avg, 4), "GBs": round(b / (avg * 1e-3) / 1e9, 1) if b and avg > 0 else None}
    byte_phases = [p for p in per_phase if pbytes(p) > 0]
    dom = max(byte_phases, key=lambda p: per_phase[p]["ms"])
    dom_bytes = pbytes(dom)
    dom_s = per_phase[dom]["ms"] * 1e-3
    achieved = dom_bytes / dom_s / 1e9
    step_b = roofline.step_bytes(R, n_avg, nc, cand, coll, n_bgk, **kw)
    pm = PMC_TRAFFIC.get(args.workload, {}) if (args.scale == 1.0 and args.precision == "fp32") else {}
    traffic = int(pm[dom]) if dom in pm else None
    traffic_src = (pm["profile"] + " (rocprofv3 --pmc FETCH_SIZE x2 + WRITE_SIZE)") if traffic else None
    step_gbs = step_b / (elapsed / args.steps) / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic",
        "config": {"workload": desc, "form": "plain (one case)", "precision":
                   "FP32 particle state, FP64 samplers/sigma_max/NTC" if R == 4 else "all FP64",
                   "state": (f"timed after {args.develop + args.warmup} steps from the "
                             + ("isentropic start" if args.workload == "expansion" else "fill")),
                   "decomposition_interval": spec.decomposition_interval},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": roofline.HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / roofline.HBM_PEAK_GBS, 4),
                     "traffic": traffic, "traffic_source": traffic_src, "bytes_per_launch": int(dom_bytes)},
        "step_bandwidth": {"algorithmic_bytes_per_step": int(step_b), "GBs": round(step_gbs, 1),
                           "frac": round(step_gbs / roofline.HBM_PEAK_GBS, 4)},
        "phases": per_phase,
        "counters_per_step": {"simulators": n_avg, "candidates": cand, "collisions": coll,
                              "bgk_redrawn": n_bgk, "dsmc_cells": c1["n_dsmc_cells"]},
        "cell_tile": case.tile_config(),
    }
    if args.probe:
        nbytes = 2 << 30
        s = probe_copy(nbytes, reps=10)
        out["hbm_probe"] = {"copy_bytes": nbytes, "GBs": round(2 * nbytes / s / 1e9, 1)}
    case.close()
    if args.precision == "fp32" and not args.no_fp64_line:
        out["fp64_line"] = fp64_line(args, spec, init)
    if args.workload == "expansion" and not args.no_nozzle_line:
        out["mixed_nozzle_line"] = nozzle_line(args)
    if args.sweep_cases > 0:
        out["sweep"] = sweep_leg(args.sweep_cases, args.sweep_rate, args.precision, args.sweep_saturated_rate)
    if not args.no_precision:
        print("[bench] FP32 vs FP64 fields ...", file=sys.stderr, flush=True)
        out["fp32_vs_fp64"] = precision_leg()
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_seconds, args.scale)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
