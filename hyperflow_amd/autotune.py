"""Tile-shape autotuner for the update kernels.

A candidate is the number of rows a wave sweeps per work item (`tile_rows`);
the strip width is fixed by the 64-lane wave.  Every candidate produces
bit-identical results by construction (the row split only changes which wave
computes a cell), and the tuner checks that: a candidate is accepted only if
a short run from the same state gives exactly the fields of the reference
candidate.  The fastest accepted candidate is stored in a JSON cache keyed by
the workload's signature, so later runs (and bench.py) reuse it.
"""
from __future__ import annotations

import hashlib
import json
import os
import time
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import native

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
CANDIDATES = (8, 16, 24, 32, 40, 42, 44, 48, 56, 64, 96)


def signature(cases, precision: str, dt_mode: str) -> str:
    c0 = cases[0]
    blocks = sorted((b.nx, b.ny) for c in cases for b in c.blocks)
    h = hashlib.sha1(json.dumps(blocks).encode()).hexdigest()[:12]
    return f"{c0.scheme}/{c0.limiter}/{precision}/{dt_mode}/p{len(cases)}/b{len(blocks)}/{h}"


def load_cache() -> Dict[str, dict]:
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def cached_rows(cases, precision: str = "fp64", dt_mode: str = "compat") -> int:
    """Tuned tile_rows for this workload, or 0 (library default)."""
    if not isinstance(cases, (list, tuple)):
        cases = [cases]
    entry = load_cache().get(signature(cases, precision, dt_mode))
    return int(entry["tile_rows"]) if entry else 0


def _time(cases, fields, precision, dt_mode, rows, steps, warmup, graph=True):
    s = native.Solver(cases, precision=precision, dt_mode=dt_mode, tile_rows=rows)
    s.upload_all(fields)
    s.run(warmup, graph=graph)
    s.synchronize()
    t = time.perf_counter()
    s.run(steps, graph=graph)
    s.synchronize()
    el = time.perf_counter() - t
    out = s.download_all()
    s.close()
    return el, out


def tune(cases, fields: Sequence[np.ndarray], precision: str = "fp64", dt_mode: str = "compat",
         candidates: Sequence[int] = CANDIDATES, steps: int = 20, warmup: int = 3,
         save: bool = True, verbose: bool = False, graph: bool = True) -> dict:
    """Times every candidate on `steps` steps from the given fields (HIP-graph
    captured, or eager launches as the workload's benchmark runs them) and
    returns {"tile_rows", "ms_per_step", "timings", "rejected"}."""
    if not isinstance(cases, (list, tuple)):
        cases = [cases]
    ref_rows = candidates[0]
    ref_time, ref_out = _time(cases, fields, precision, dt_mode, ref_rows, steps, warmup, graph)
    timings = {ref_rows: ref_time / steps * 1e3}
    rejected: List[int] = []
    for rows in candidates[1:]:
        el, out = _time(cases, fields, precision, dt_mode, rows, steps, warmup, graph)
        same = all(np.array_equal(a, b, equal_nan=True) for a, b in zip(out, ref_out))
        if not same:
            rejected.append(rows)
            continue
        timings[rows] = el / steps * 1e3
        if verbose:
            print(f"tile_rows {rows}: {timings[rows]:.4f} ms/step", flush=True)
    best = min(timings, key=timings.get)
    result = {"tile_rows": int(best), "ms_per_step": timings[best],
              "timings": {str(k): v for k, v in sorted(timings.items())}, "rejected": rejected}
    if save:
        cache = load_cache()
        cache[signature(cases, precision, dt_mode)] = result
        with open(CACHE, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return result


def _workloads(names):
    """The benchmark workloads by name: (label, cases, fields, precision,
    dt_mode, graph), graph as bench.py launches the workload."""
    from . import case as C
    out = []
    for name in names:
        if name == "godunov-quadrants":
            c = C.godunov_quadrants(0)
            out.append(("godunov-quadrants", [c], [c.initial_fields()[0]], "fp64", "compat", False))
        elif name == "step-channel":
            c = C.muscl_step_channel(1024)
            out.append(("step-channel", [c], c.initial_fields(), "fp64", "compat", False))
        elif name in ("muscl-quadrants", "muscl-quadrants-fp32"):
            c = C.muscl_quadrants(0, 8192)
            f = C.perturbed_fields(c, 0)
            for prec in (("fp64", "fp32") if name == "muscl-quadrants" else ("fp32",)):
                out.append((f"muscl-quadrants-{prec}", [c], f, prec, "compat", False))
        elif name == "batch":
            cases = [C.quad_case(nx, ny, 1000 + k, scheme="MUSCLHancock", limiter="VanLeer", cfl=0.5)
                     for k, (nx, ny) in enumerate(C.batched_sizes(512, 0))]
            out.append(("batch", cases, [c.initial_fields()[0] for c in cases], "fp32", "compat", True))
        elif name == "learned":
            for n in (1024, 1152, 1280, 1536, 1792, 2048):
                c = C.muscl_quadrants(0, n)
                out.append((f"learned-{n}", [c], C.perturbed_fields(c, 0), "fp32", "uniform", True))
        else:
            raise SystemExit(f"unknown workload {name}")
    return out


def main(argv=None) -> int:
    import argparse
    ap = argparse.ArgumentParser(prog="python -m hyperflow_amd.autotune",
                                 description="tune tile rows of the benchmark workloads into the cache")
    ap.add_argument("workloads", nargs="*", default=["godunov-quadrants", "step-channel", "muscl-quadrants", "batch", "learned"])
    ap.add_argument("--steps", type=int, default=20, help="timed steps per candidate")
    args = ap.parse_args(argv)
    for label, cases, fields, prec, mode, graph in _workloads(args.workloads):
        r = tune(cases, fields, precision=prec, dt_mode=mode, steps=args.steps, graph=graph)
        print(f"{label}: tile_rows {r['tile_rows']} ({r['ms_per_step']:.4f} ms/step) {r['timings']}", flush=True)
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(main())
