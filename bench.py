"""Benchmark: MUSCL-Hancock + MinBee + HLLC, FP64, on the 8-block step channel
(eight 1024x1024 blocks, Mach 3 inflow, reflective walls and step, compat dt),
synthetic initial data (the inflow state), one MI355X.

    python bench.py [--gpus 1] [--steps 500] [--warmup 20]

A step is one full time step of the path: ghost fill, CFL timestep reduction
and the eight fused per-block MUSCL-HLLC updates.  The timed region starts
with the state resident in HBM.  Prints one JSON line; its `learned`
sub-object is the learned correction's speed-up at equal density L1 on one
unseen seed (`--workload learned` below, about a minute, in a child process;
--no-learned leaves it out).

Other workloads of the same path (not the headline line):
    --workload godunov-quadrants
        Godunov + HLLC, 400x400 FP64, transmissive, CFL 0.8
    --workload muscl-quadrants
        MUSCL-Hancock + Van Leer, 8192x8192 (--precision fp32|fp64)
    --workload batch
        512 independent problems of U[64,512]^2 cells, MUSCL FP32, one
        persistent launch per step, HIP graph; also times looping the same
        problems one context at a time
    --workload learned
        learned coarse correction: 1024^2 MUSCL FP32 + the committed CNN
        against plain runs at 1024..1792^2 and the 2048^2 solution (density
        L1, wall time, speed-up at equal L1)
    --workload copy
        a 1 GiB device copy: measured copy bandwidth, and the calibration run
        for the HBM traffic counters
    --workload unstructured
        unstructured Godunov (fvm_2d_uns), FP64: the shipped scramjet case on
        its coarse block mesh refined --refine times per direction (default 8:
        1 469 440 cells)
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
# the update launches of every 10th timed step carry HIP timing events (stamped by the launch itself):
# the average launch duration is sampled over the whole timed region at a tenth of the timing cost
PROFILE_EVERY = 10
# The FP64 MUSCL-Hancock row loop of k_muscl (one row of 60 output cells per
# wave), counted in the gfx950 assembly: MinBee 1 397 VALU instructions, Van
# Leer 1 429, 60 of them quarter-rate (v_rcp_f64 / v_rsq_f64).  A wave64 VALU
# instruction holds its SIMD for 4 cycles (16 for quarter rate); 1 024 SIMDs
# at 2.4 GHz.
VALU_ROW = {("MUSCLHancock", "MinBee", "fp64"): (1397, 60), ("MUSCLHancock", "VanLeer", "fp64"): (1429, 60)}
VALU_SIMDS, VALU_GHZ, VALU_CELLS_PER_ROW = 1024, 2.4, 60


def valu_bound(scheme, limiter, precision, cells_per_s):
    """The update kernel's cell rate against its VALU issue bound (None where
    the row loop has not been counted)."""
    key = (scheme, limiter, precision)
    if key not in VALU_ROW:
        return None
    n, quarter = VALU_ROW[key]
    cycles = (n - quarter) * 4 + quarter * 16
    bound = VALU_SIMDS * VALU_GHZ * 1e9 / cycles * VALU_CELLS_PER_ROW / 1e6
    achieved = cells_per_s / 1e6
    return {"unit": "Mcell-updates/s", "achieved": round(achieved, 1), "bound": round(bound, 1),
            "frac": round(achieved / bound, 4), "valu_per_row": n, "quarter_rate_per_row": quarter,
            "cells_per_row": VALU_CELLS_PER_ROW}


def cpu_baseline(sample_n: int = 256, sample_steps: int = 5):
    """Times the reference's own FP64 CPU path (oracle/_ref/ref_harness, built
    from the reference sources with its makefile's flags: serial, one core) on
    a bounded sample of the same workload.  Falls back to the CPU oracle."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import hfo
    from hyperflow_amd import case as C

    case = C.muscl_step_channel(sample_n)
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


def copy_bandwidth(nbytes: int = 1 << 30, reps: int = 20):
    """A plain device copy of a known size (read nbytes + write nbytes per
    launch), timed with HIP events: the measured HBM copy bandwidth, and the
    calibration kernel for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    a = torch.full((nbytes // 4,), 1.0, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        b.copy_(a)
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / reps
    return {"metric": "device copy bandwidth", "value": round(2 * nbytes / (ms * 1e-3) / 1e9, 1), "unit": "GB/s",
            "bytes_read_per_launch": nbytes, "bytes_written_per_launch": nbytes, "avg_launch_ms": round(ms, 4),
            "launches": reps + 3}


LEARNED_RESOLUTIONS = (1024, 1152, 1280, 1536, 1792)
# Unseen seeds (training used 0-39).  The shipped network and its settings were
# chosen on 5000-5002, 6000-6002 and 7000-7002; 8000-8002 took no part in any
# choice and give the headline sub-line.
LEARNED_SEEDS_CHOICE = (5000, 5001, 5002, 6000, 6001, 6002, 7000, 7001, 7002)
LEARNED_SEEDS_CHECK = (8000, 8001, 8002)


def learned_line(seeds=LEARNED_SEEDS_CHOICE + LEARNED_SEEDS_CHECK, coarse_steps: int = 200,
                 resolutions=LEARNED_RESOLUTIONS):
    """The learned-correction comparison on unseen seeds: density L1 against
    the 2x2-averaged 2048^2 solution, and wall times (HIP graphs, state in
    HBM), for plain runs at the given resolutions and the corrected 1024^2
    run."""
    from hyperflow_amd import learned as L
    L.ensure_tuned(tuple(resolutions) + (2048,), log=lambda m: print(m, file=sys.stderr))
    net = L.load()
    meta = L.load_meta()
    every, start = int(meta["every"]), int(meta["start"])
    rows, summ = L.evaluate(net, list(seeds), coarse_steps, resolutions=tuple(resolutions), every=(every,),
                            start=start, log=lambda m: print(m, file=sys.stderr))
    lr = summ["learned"][every]
    return {"metric": "learned vs plain time at equal L1", "value": round(lr["speedup_at_equal_l1"], 4), "unit": "x",
            "higher_is_better": True, "n_gpus": 1, "dtype": "f32 (solver), bf16/f32 (network)",
            "data": "synthetic (seeded quadrant states with sinusoidal density perturbation)",
            "config": {"workload": "MUSCL-Hancock+VanLeer+HLLC FP32 on [0,1]^2, transmissive, fixed Courant 0.4/6.6; "
                                   f"learned 1024^2 corrected every {every} steps from step {start} vs plain n^2",
                       "coarse_steps": coarse_steps, "seeds": list(seeds)},
            "learned_l1": lr["l1"], "learned_seconds": lr["seconds"], "plain_resolutions": summ["resolutions"],
            "plain_l1": summ["plain_l1"], "plain_seconds": summ["plain_seconds"],
            "plain_2048_seconds": summ["fine_seconds"], "plain_n_at_equal_l1": lr["plain_n_at_equal_l1"],
            "plain_seconds_at_equal_l1": lr["plain_seconds_at_equal_l1"],
            "speedup_at_equal_l1_n3_fit": lr["speedup_at_equal_l1_fit"],
            "l1_gain_vs_plain_1024": lr["l1_gain_vs_plain_1024"], "speedup_vs_plain_2048": lr["speedup_vs_plain_2048"],
            "l1_gain_per_seed": {str(r["seed"]): r["plain"][resolutions[0]]["l1"] / r["learned"][every]["l1"]
                                 for r in rows},
            "vs_baseline": None}


def learned_subline(timeout_s: int = 240):
    """The learned comparison for the headline line, on the three seeds that
    took no part in choosing the network (about a minute): run as a child process, because the solver library and torch
    each bring a HIP runtime and a process can use only the one it loaded
    first."""
    cmd = [sys.executable, os.path.abspath(__file__), "--workload", "learned", "--learned-quick"]
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout_s)
    except subprocess.TimeoutExpired:
        return {"error": f"timed out after {timeout_s} s"}
    if r.returncode != 0:
        return {"error": f"exit {r.returncode}: " + r.stderr.strip().splitlines()[-1][:300] if r.stderr.strip() else
                f"exit {r.returncode}"}
    d = json.loads(r.stdout.strip().splitlines()[-1])
    keep = ("value", "unit", "learned_l1", "learned_seconds", "plain_resolutions", "plain_l1", "plain_seconds",
            "plain_n_at_equal_l1", "plain_seconds_at_equal_l1", "l1_gain_vs_plain_1024", "speedup_vs_plain_2048",
            "l1_gain_per_seed")
    out = {"metric": "learned vs plain time at equal L1 (speed-up)"}
    out.update({k: d[k] for k in keep if k in d})
    out["config"] = d.get("config")
    return out


def traffic_from_profiles(workload: str):
    """HBM bytes per update launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE in separate runs, each scaled by what the counter reads for a
    copy of known size), committed in profiles/hbm_traffic.json."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get(workload, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def uns_line(steps: int, warmup: int, refine: int, with_cpu: bool):
    """The unstructured solver on the shipped scramjet case: its coarse block
    mesh (the reference's scramjet_kumar_coarse_mesh.json, kept with the
    scramjet golden) with every block's cell counts multiplied by `refine`,
    Constant inflow initial state, device-resident steps (one launch each)."""
    import copy
    from hyperflow_amd import uns as U
    with open(os.path.join(ROOT, "tests", "golden", "uns_kumar_coarse.json")) as f:
        spec = json.load(f)
    blocks = copy.deepcopy(spec["mesh_blocks"])
    for b in blocks:
        b["cells"] = [b["cells"][0] * refine, b["cells"][1] * refine]
    q, c = U.block_mesh(blocks)
    case = spec["case"]
    gamma, cfl = case["model"]["gamma"], case["simulation"]["cfl"]
    s = U.UnsSolver(q, c, U.prim_to_cons(case["mesh"]["boundaries"]["inlet"], gamma), gamma, cfl)
    s.upload(U.initial_fields(q, case["initial_condition"], gamma))
    bpc = s.bytes_per_cell()   # state in + out, neighbour indices, lengths, rotation terms, 1/V
    s.run(warmup)
    s.synchronize()
    s.profile(True)
    s.kernel_stats()
    t0 = time.perf_counter()
    s.run(steps)
    s.synchronize()
    wall = time.perf_counter() - t0
    launches, kernel_ms = s.kernel_stats()
    s.profile(False)
    s.run(steps, graph=True)       # capture, then time the same loop as a HIP graph
    s.synchronize()
    t0 = time.perf_counter()
    s.run(steps, graph=True)
    s.synchronize()
    wall_graph = time.perf_counter() - t0
    s.close()
    n = len(q)
    avg_ms = kernel_ms / launches
    achieved = n * bpc / (avg_ms * 1e-3) / 1e9
    line = {
        "metric": "unstructured Godunov-HLLC Mcell-updates/s", "value": round(n * steps / wall / 1e6, 2),
        "unit": "Mcell-updates/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": round(wall / steps * 1e3, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (the case's constant inflow state)",
        "config": {"workload": f"scramjet_kumar_godunov_hllc on the coarse block mesh refined {refine}x per "
                               f"direction ({n} quadrilateral cells), Godunov + rotated HLLC, forward Euler",
                   "cells": n, "cfl": cfl},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(f"unstructured_r{refine}"),
                     "kernel": "k_uns_step", "avg_launch_ms": round(avg_ms, 5), "launches": launches,
                     "algorithmic_bytes_per_launch": n * bpc, "bytes_per_cell": bpc},
        "hip_graph": {"ms_per_step": round(wall_graph / steps * 1e3, 5),
                      "value": round(n * steps / wall_graph / 1e6, 2),
                      "note": "the same steps captured in a HIP graph (one replay, after the capturing run)"},
    }
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import hfo
        q0, c0 = U.block_mesh(spec["mesh_blocks"])
        o = hfo.UnsOracle(q0, c0, U.prim_to_cons(case["mesh"]["boundaries"]["inlet"], gamma), gamma, cfl)
        o.set_fields(U.initial_fields(q0, case["initial_condition"], gamma))
        k = 400
        t = time.perf_counter()
        o.run(k)
        el = time.perf_counter() - t
        line["cpu_baseline"] = {"value": round(len(q0) * k / el / 1e6, 3), "unit": "Mcell-updates/s", "cores": 1,
                                "kind": "port",
                                "sample": f"the same case on the unrefined coarse mesh ({len(q0)} cells), {k} steps, "
                                          "CPU oracle (C, serial); the reference itself ran 0.075 Mcell-updates/s "
                                          "on this mesh when its goldens were written"}
    return line


def build_workload(name: str, n: int, precision: str, seed: int = 0):
    from hyperflow_amd import case as C
    if name == "godunov-quadrants":
        case = C.godunov_quadrants(seed)
        return [case], [case.initial_fields()[0]], "Godunov+HLLC FP64, 400x400, transmissive, CFL 0.8"
    if name == "step-channel":
        case = C.muscl_step_channel(n)
        return [case], case.initial_fields(), (
            f"MUSCL-Hancock+MinBee+HLLC {precision.upper()}, 8 blocks of {n}x{n} step channel (Dirichlet inflow, "
            "reflective walls, transmissive outflow), CFL 0.5, reference-compatible dt")
    if name == "muscl-quadrants":
        from hyperflow_amd.accuracy import C3_SEED
        n3 = 8192 if n == 1024 else n
        case = C.muscl_quadrants(C3_SEED, n3)
        return [case], C.perturbed_fields(case, C3_SEED), (
            f"MUSCL-Hancock+VanLeer+HLLC {precision.upper()}, single block {n3}x{n3}, seeded quadrants with "
            f"sinusoidal density perturbation (seed {C3_SEED}), CFL 0.5")
    if name == "batch":
        cases = [C.quad_case(nx, ny, 1000 + k, scheme="MUSCLHancock", limiter="VanLeer", cfl=0.5)
                 for k, (nx, ny) in enumerate(C.batched_sizes(512, seed))]
        return cases, [c.initial_fields()[0] for c in cases], (
            f"512 independent seeded 2D Riemann problems, sides U[64,512], MUSCL-HLLC {precision.upper()}, "
            "one persistent tile-queue launch per step, HIP-graph captured")
    raise SystemExit(f"unknown workload {name}")


def looped_baseline(cases, fields, precision, steps):
    """Each problem in its own context, advanced one after another."""
    from hyperflow_amd import native
    solvers = []
    for c, f in zip(cases, fields):
        s = native.Solver(c, precision=precision)
        s.upload(0, f)
        solvers.append(s)
    for s in solvers:
        s.run(2)
    for s in solvers:
        s.synchronize()
    t = time.perf_counter()
    for s in solvers:
        s.run(steps)
    for s in solvers:
        s.synchronize()
    el = time.perf_counter() - t
    for s in solvers:
        s.close()
    return el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", choices=["godunov-quadrants", "step-channel", "muscl-quadrants", "batch", "learned", "copy", "unstructured"], default="step-channel")
    ap.add_argument("--refine", type=int, default=8, help="unstructured: cells per coarse cell and direction")
    ap.add_argument("--n", type=int, default=1024, help="cells per block side (step-channel), grid side (muscl-quadrants)")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default=None)
    ap.add_argument("--tile-rows", type=int, default=None, help="default: the autotuned work plan of the class")
    ap.add_argument("--graph", action="store_true", help="capture the timed steps in a HIP graph")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true",
                    help="leave the learned comparison out of the step-channel line")
    ap.add_argument("--learned-quick", action="store_true",
                    help="learned: the three check seeds and plain runs up to 1536^2 (the headline line's sub-object)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the path would shard by blocks with a halo exchange and a dt min-allreduce; not implemented)",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "copy":
        print(json.dumps(copy_bandwidth()), flush=True)
        return
    if args.workload == "learned":
        if args.learned_quick:
            print(json.dumps(learned_line(seeds=LEARNED_SEEDS_CHECK, resolutions=(1024, 1152, 1280, 1536))), flush=True)
        else:
            print(json.dumps(learned_line()), flush=True)
        return
    if args.workload == "unstructured":
        print(json.dumps(uns_line(args.steps, args.warmup, args.refine, not args.no_cpu_baseline)), flush=True)
        return

    from hyperflow_amd import autotune, native

    precision = args.precision or ("fp32" if args.workload == "batch" else "fp64")
    cases, fields, workload_text = build_workload(args.workload, args.n, precision)
    plan = autotune.cached_params(cases, precision)
    if args.tile_rows is not None:
        plan = {"tile_rows": args.tile_rows, "tail_rows": 0, "fix_rows": 0}
    graph = args.graph or args.workload == "batch"

    solver = native.Solver(cases, precision=precision, **plan)
    solver.upload_all(fields)
    solver.run(args.warmup, graph=graph)
    solver.synchronize()

    solver.profile(not graph, every=PROFILE_EVERY)
    solver.kernel_stats()                         # reset counters
    t0 = time.perf_counter()
    solver.run(args.steps, graph=graph)
    solver.synchronize()
    t1 = time.perf_counter()
    launches, kernel_ms = solver.kernel_stats()
    solver.profile(False)
    solver.close()          # release the context before anything else runs (no teardown at interpreter exit)

    wall = t1 - t0
    cells = sum(c.interior_cells for c in cases)
    value = cells * args.steps / wall / 1e6
    bpc = BYTES_PER_CELL[precision]
    if launches:
        avg_ms = kernel_ms / launches
        timed_steps = (args.steps + PROFILE_EVERY - 1) // PROFILE_EVERY   # steps 0, P, 2P, ... of the region
        cells_per_launch = cells * timed_steps / launches
    else:                                         # graph mode: no per-launch events
        avg_ms = wall * 1e3 / args.steps
        cells_per_launch = cells
    alg_bytes = cells_per_launch * bpc
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    # counters were taken at the default sizes (1024 per block side, 8192 for muscl-quadrants)
    traffic = traffic_from_profiles(f"{args.workload}_{precision}") if args.n == 1024 else None
    if cases[0].scheme != "MUSCLHancock":
        kernel = "k_godunov"
    elif len(cases) == 1 and len(cases[0].blocks) > 1:
        kernel = ("the step's MUSCL update launches k_muscl (main list), k_muscl_guess and k_muscl_check, "
                  "timed from the first one's start to the last one's end")
    else:
        kernel = "k_muscl"

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
        "dtype": "f64" if precision == "fp64" else "f32",
        "data": ("synthetic (the channel filled with its Mach 3 inflow state, as in the reference case)"
                 if args.workload == "step-channel" else "synthetic (seeded quadrant states of the named workload)"),
        "config": {"workload": workload_text, "blocks": sum(len(c.blocks) for c in cases),
                   "interior_cells": cells, "scheme": cases[0].scheme, "limiter": cases[0].limiter,
                   "riemann": "HLLC", "cfl": cases[0].cfl,
                   "work_plan": {k: (v or "default") for k, v in plan.items()}, "hip_graph": graph},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": kernel if launches else "whole step (graph)", "avg_launch_ms": round(avg_ms, 5),
                     "timed_every_nth_step": PROFILE_EVERY if launches else None,
                     "launches": launches, "algorithmic_bytes_per_launch": int(alg_bytes),
                     "bytes_per_cell": bpc},
    }
    if launches:
        valu = valu_bound(cases[0].scheme, cases[0].limiter, precision, cells_per_launch / (avg_ms * 1e-3))
        if valu:
            line["roofline"]["valu"] = valu
    if args.workload == "batch":
        loop = looped_baseline(cases, fields, precision, args.steps)
        line["looped_per_problem_s"] = round(loop, 4)
        line["batched_speedup_vs_looped"] = round(loop / wall, 2)
    if not args.no_cpu_baseline and args.workload == "step-channel":
        line["cpu_baseline"] = cpu_baseline()
    if not args.no_learned and args.workload == "step-channel":
        line["learned"] = learned_subline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
