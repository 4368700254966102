"""Work-plan autotuner for the update kernels.

A work plan is how a step's update launch cuts the blocks into work items:
`tile_rows` (rows a wave sweeps per item of the main list), `tail_rows` (rows
per item of the main list's closing phase, sized so the last round of items
ends together) and `fix_rows` (rows per item of the fix-up rounds of the
reference-compatible predictor dt).  The strip width is fixed by the 64-lane
wave.  No plan changes a result: the split only decides which wave computes
a cell.  The tuner checks that anyway: a candidate is accepted only if a
short run from the same state gives exactly the fields of the default plan.

Plans are stored per problem class, not per exact workload: scheme, limiter,
precision, dt mode, the number of problems (1, 2-63, 64 or more), blocks per
problem (1, 2-8, 9 or more) and the median block size (cells, to the nearest
power of 4).  A workload whose class has no entry gets the plan of the
nearest tuned class with the same scheme, limiter, precision and dt mode.

Every candidate is timed `repeats` times (each run from the same state, after
warm-up steps); its time is the median, and the chosen plan records the
spread (max - min) / median of its runs.

    python -m hyperflow_amd.autotune [workload ...]
"""
from __future__ import annotations

import hashlib
import json
import math
import os
import time
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import native

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
ROWS = (8, 16, 24, 32, 40, 42, 44, 48, 56, 64)
TAILS = (4, 6, 8, 12, 16)
FIXES = (0, 4, 6, 9, 12, 16)   # 0: sized by the library to one round of the resident waves
PARAMS = ("tile_rows", "tail_rows", "fix_rows")


def _cases(cases):
    return list(cases) if isinstance(cases, (list, tuple)) else [cases]


def problem_class(cases, precision: str = "fp64", dt_mode: str = "compat") -> dict:
    """The class a workload is tuned and looked up under."""
    cases = _cases(cases)
    c0 = cases[0]
    np_ = len(cases)
    bpp = max(len(c.blocks) for c in cases)
    cells = sorted(b.nx * b.ny for c in cases for b in c.blocks)
    med = cells[len(cells) // 2]
    return {"scheme": c0.scheme, "limiter": c0.limiter, "precision": precision, "dt_mode": dt_mode,
            "problems": "1" if np_ == 1 else ("2-63" if np_ < 64 else "64+"),
            "blocks": "1" if bpp == 1 else ("2-8" if bpp <= 8 else "9+"),
            "size4": int(round(math.log(max(med, 1), 4)))}


def class_key(cls: dict) -> str:
    return (f"{cls['scheme']}/{cls['limiter']}/{cls['precision']}/{cls['dt_mode']}/p{cls['problems']}"
            f"/b{cls['blocks']}/n4^{cls['size4']}")


def load_cache() -> Dict[str, dict]:
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def lookup(cases, precision: str = "fp64", dt_mode: str = "compat") -> Optional[dict]:
    """The cache entry of the workload's class, else of the nearest tuned class
    with the same scheme, limiter, precision and dt mode (block size counts one
    per power of 4, a different problem-count or blocks-per-problem bucket
    three), else None."""
    cls = problem_class(cases, precision, dt_mode)
    cache = load_cache()
    exact = cache.get(class_key(cls))
    if exact:
        return exact
    best, best_d = None, None
    for e in cache.values():
        k = e.get("class")
        if not k or any(k[f] != cls[f] for f in ("scheme", "limiter", "precision", "dt_mode")):
            continue
        d = abs(k["size4"] - cls["size4"]) + 3 * (k["problems"] != cls["problems"]) + 3 * (k["blocks"] != cls["blocks"])
        if best_d is None or d < best_d:
            best, best_d = e, d
    return best


def has_class(cases, precision: str = "fp64", dt_mode: str = "compat") -> bool:
    """Whether the workload's own class has been tuned."""
    return class_key(problem_class(cases, precision, dt_mode)) in load_cache()


def cached_params(cases, precision: str = "fp64", dt_mode: str = "compat") -> dict:
    """Tuned work plan {tile_rows, tail_rows, fix_rows} for the workload's
    class (zeros: the library defaults)."""
    e = lookup(cases, precision, dt_mode)
    return {p: int(e[p]) if e else 0 for p in PARAMS}


def cached_rows(cases, precision: str = "fp64", dt_mode: str = "compat") -> int:
    """Tuned tile_rows for the workload's class, or 0 (library default)."""
    return cached_params(cases, precision, dt_mode)["tile_rows"]


def _digest(fields: Sequence[np.ndarray]) -> str:
    h = hashlib.sha256()
    for U in fields:
        h.update(np.ascontiguousarray(U).tobytes())
    return h.hexdigest()


def _run(cases, fields, precision, dt_mode, plan, steps, warmup, graph):
    s = native.Solver(cases, precision=precision, dt_mode=dt_mode, **plan)
    try:
        s.upload_all(fields)
        s.run(warmup, graph=graph)
        s.synchronize()
        t = time.perf_counter()
        s.run(steps, graph=graph)
        s.synchronize()
        el = time.perf_counter() - t
        digest = _digest(s.download_all())
    finally:
        s.close()
    return el / steps * 1e3, digest


def tune(cases, fields: Sequence[np.ndarray], precision: str = "fp64", dt_mode: str = "compat",
         rows: Sequence[int] = ROWS, tails: Sequence[int] = TAILS, fixes: Sequence[int] = FIXES,
         steps: int = 20, warmup: int = 3, repeats: int = 3, save: bool = True, verbose: bool = False,
         graph: bool = True, label: str = "") -> dict:
    """Coordinate descent over tile_rows, then tail_rows, then fix_rows (only
    for compat MUSCL problems of several blocks), each candidate timed
    `repeats` times on `steps` steps (HIP-graph captured, or eager launches as
    the workload's benchmark runs them) from the given fields.  Returns the
    chosen plan with its median time and spread, every candidate's times and
    the candidates rejected for differing results."""
    cases = _cases(cases)
    default = {p: 0 for p in PARAMS}
    _, ref = _run(cases, fields, precision, dt_mode, default, steps, warmup, graph)
    timings: Dict[str, List[float]] = {}
    rejected: List[str] = []

    def measure(plan) -> Optional[float]:
        key = ",".join(str(plan[p]) for p in PARAMS)
        if key in timings:
            return float(np.median(timings[key]))
        if key in rejected:
            return None
        runs = []
        for _ in range(repeats):
            ms, d = _run(cases, fields, precision, dt_mode, plan, steps, warmup, graph)
            if d != ref:   # the bitwise gate: the same fields as the default plan after the same steps
                rejected.append(key)
                return None
            runs.append(ms)
        timings[key] = runs
        if verbose:
            print(f"  {key}: {np.median(runs):.4f} ms/step ({', '.join(f'{r:.4f}' for r in runs)})", flush=True)
        return float(np.median(runs))

    best = dict(default)
    best["tile_rows"] = rows[0]
    best_t = measure(best)
    for p, cands in (("tile_rows", rows), ("tail_rows", tails), ("fix_rows", fixes)):
        if p == "fix_rows" and not (dt_mode == "compat" and cases[0].scheme == "MUSCLHancock"
                                    and max(len(c.blocks) for c in cases) > 1):
            continue
        for v in cands:
            plan = dict(best)
            plan[p] = v
            t = measure(plan)
            if t is not None and (best_t is None or t < best_t):
                best, best_t = plan, t
    if best_t is None:
        raise RuntimeError("autotune: no candidate reproduced the default plan's fields")
    key = ",".join(str(best[p]) for p in PARAMS)
    runs = timings[key]
    cls = problem_class(cases, precision, dt_mode)
    result = dict(best)
    result.update({"ms_per_step": best_t, "spread": (max(runs) - min(runs)) / best_t, "repeats": repeats,
                   "timings": {k: [round(x, 5) for x in v] for k, v in timings.items()}, "rejected": rejected,
                   "class": cls, "workload": label})
    if save:
        cache = load_cache()
        cache[class_key(cls)] = result
        with open(CACHE, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return result


def _workloads(names):
    """The benchmark workloads by name: (label, cases, fields, precision,
    dt_mode, graph, search), graph as bench.py launches the workload, search
    the candidate lists and repeats (fewer for the 8192^2 block, whose every
    run moves 4 GB between host and device)."""
    from . import case as C
    out = []
    for name in names:
        if name == "godunov-quadrants":
            c = C.godunov_quadrants(0)
            # 160k cells fill a fraction of the resident waves: short items
            # keep more SIMDs busy (a Godunov item has one halo row per side)
            out.append(("godunov-quadrants", [c], [c.initial_fields()[0]], "fp64", "compat", False,
                        {"rows": (2, 3, 4, 6, 8, 12, 16), "tails": (1, 2, 3, 4, 6, 8)}))
        elif name == "step-channel":
            c = C.muscl_step_channel(1024)
            # past the first 10 TimeStep calls' reduced CFL, as the benchmark times it
            out.append(("step-channel", [c], c.initial_fields(), "fp64", "compat", False, {"warmup": 20}))
        elif name in ("muscl-quadrants", "muscl-quadrants-fp32"):
            from .accuracy import C3_SEED
            c = C.muscl_quadrants(C3_SEED, 8192)
            f = C.perturbed_fields(c, C3_SEED)
            for prec in (("fp64", "fp32") if name == "muscl-quadrants" else ("fp32",)):
                out.append((f"muscl-quadrants-{prec}", [c], f, prec, "compat", False,
                            {"rows": (32, 48, 64, 96), "tails": (4, 8, 16), "repeats": 2}))
        elif name == "batch":
            cases = [C.quad_case(nx, ny, 1000 + k, scheme="MUSCLHancock", limiter="VanLeer", cfl=0.5)
                     for k, (nx, ny) in enumerate(C.batched_sizes(512, 0))]
            out.append(("batch", cases, [c.initial_fields()[0] for c in cases], "fp32", "compat", True, {}))
        elif name == "learned":
            for n in (1024, 1536, 2048):
                c = C.muscl_quadrants(0, n)
                out.append((f"learned-{n}", [c], C.perturbed_fields(c, 0), "fp32", "uniform", True,
                            {"repeats": 2}))
        else:
            raise SystemExit(f"unknown workload {name}")
    return out


def main(argv=None) -> int:
    import argparse
    ap = argparse.ArgumentParser(prog="python -m hyperflow_amd.autotune",
                                 description="tune the work plan of the benchmark workloads into the cache")
    ap.add_argument("workloads", nargs="*",
                    default=["godunov-quadrants", "step-channel", "muscl-quadrants", "batch", "learned"])
    ap.add_argument("--steps", type=int, default=20, help="timed steps per run")
    ap.add_argument("--repeats", type=int, default=3, help="timed runs per candidate")
    args = ap.parse_args(argv)
    for label, cases, fields, prec, mode, graph, search in _workloads(args.workloads):
        print(f"{label}: class {class_key(problem_class(cases, prec, mode))}", flush=True)
        search = dict(search)
        repeats = search.pop("repeats", args.repeats)
        search.setdefault("warmup", 3)
        r = tune(cases, fields, precision=prec, dt_mode=mode, steps=args.steps, repeats=repeats, graph=graph,
                 label=label, verbose=True, **search)
        print(f"{label}: tile_rows {r['tile_rows']} tail_rows {r['tail_rows']} fix_rows {r['fix_rows']} "
              f"({r['ms_per_step']:.4f} ms/step, spread {100 * r['spread']:.1f} %, rejected {r['rejected']})",
              flush=True)
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(main())
