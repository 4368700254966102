"""Benchmark: MUSCL-Hancock + MinBee + HLLC, FP64, on the 8-block step channel
(eight 1024x1024 blocks, Mach 3 inflow, reflective walls and step, compat dt),
synthetic initial data (the inflow state), one MI355X.

    python bench.py [--gpus 1] [--steps 500] [--warmup 20]

A step is one full time step of the path: ghost fill, CFL timestep reduction
and the eight fused per-block MUSCL-HLLC updates.  The timed region starts
with the state resident in HBM.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MUSCL-HLLC Mcell-updates/s, % HBM roofline; learned vs plain time at equal L1"
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
BYTES_PER_CELL = {"fp64": 64, "fp32": 32}   # read U^n + write U^{n+1}, 4 conserved variables


def cpu_baseline(sample_n: int = 256, sample_steps: int = 5):
    """Times the reference's own FP64 CPU path (oracle/_ref/ref_harness, built
    from the reference sources with its makefile's flags: serial, one core) on
    a bounded sample of the same workload.  Falls back to the CPU oracle."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import hfo
    from hyperflow_amd import case as C

    case = C.config2(sample_n)
    cells = case.interior_cells * sample_steps
    sample = f"step channel 8 blocks x {sample_n}^2, MUSCL MinBee FP64, {sample_steps} steps"
    if hfo.have_reference():
        try:
            _, _, _, elapsed = hfo.run_reference(case, sample_steps, timeout=600)
            return {"value": cells / elapsed / 1e6, "unit": "Mcell-updates/s", "cores": 1, "kind": "reference",
                    "sample": sample + " (reference build, -O3, serial)"}
        except Exception as e:  # pragma: no cover - reported, not fatal
            print(f"reference baseline failed: {e}", file=sys.stderr)
    nthreads = min(16, os.cpu_count() or 1)
    hfo.set_threads(nthreads)
    m = hfo.OracleMesh(case)
    t = time.perf_counter()
    m.run(sample_steps)
    elapsed = time.perf_counter() - t
    return {"value": cells / elapsed / 1e6, "unit": "Mcell-updates/s", "cores": nthreads, "kind": "port",
            "sample": sample + " (CPU oracle, OpenMP)"}


def traffic_from_profiles(workload: str):
    """HBM bytes per update launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, calibrated against a copy kernel), committed under profiles/."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get(workload, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1024, help="cells per block side")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported by this build "
              "(the path would shard by blocks with a halo exchange and a dt min-allreduce, not measured here)",
              file=sys.stderr)
        sys.exit(2)

    from hyperflow_amd import case as C
    from hyperflow_amd import native

    case = C.config2(args.n)
    solver = native.Solver(case, precision=args.precision, tile_rows=args.tile_rows)
    solver.upload_all(case.initial_fields())
    solver.run(args.warmup)
    solver.synchronize()

    solver.profile(True)
    solver.kernel_stats()                         # reset counters
    t0 = time.perf_counter()
    solver.run(args.steps)
    solver.synchronize()
    t1 = time.perf_counter()
    launches, kernel_ms = solver.kernel_stats()
    solver.profile(False)

    wall = t1 - t0
    cells = case.interior_cells
    value = cells * args.steps / wall / 1e6
    avg_ms = kernel_ms / max(launches, 1)
    cells_per_launch = cells * args.steps / max(launches, 1)
    alg_bytes = cells_per_launch * BYTES_PER_CELL[args.precision]
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    workload = f"step_channel_{args.n}_{args.precision}"
    traffic = traffic_from_profiles(workload)

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic: uniform Mach 3 inflow state in a channel with a step",
        "config": {"workload": f"MUSCL-Hancock+MinBee+HLLC {args.precision.upper()}, 8 blocks of "
                               f"{args.n}x{args.n} step channel (Dirichlet inflow, reflective walls, "
                               f"transmissive outflow), CFL 0.5, reference-compatible dt",
                   "blocks": 8, "block_cells": [args.n, args.n], "scheme": "MUSCLHancock", "limiter": "MinBee",
                   "riemann": "HLLC", "cfl": 0.5, "dt_mode": "compat"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_muscl (per-block update)", "avg_launch_ms": round(avg_ms, 5),
                     "launches": launches, "algorithmic_bytes_per_launch": int(alg_bytes)},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
