"""Helpers shared by the GPU tests: run a case through the C ABI and through
the CPU oracle on the same inputs."""
import sys

import numpy as np

import hfo
from hyperflow_amd import native


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_run(cases, fields, nsteps, graph=False, **kw):
    s = native.Solver(cases, **kw)
    s.upload_all(fields)
    s.run(nsteps, graph=graph)
    s.synchronize()
    out = s.download_all()
    hist = s.dt_history()
    return s, out, hist


def oracle_run(case, fields, nsteps, threads=8):
    hfo.set_threads(threads)
    m = hfo.OracleMesh(case, fields)
    stepdt, calls = m.run(nsteps)
    return m.all_fields(), calls, stepdt


def rel_linf(a, b):
    """max |a-b| / max |b| over all blocks and variables (per variable scale)."""
    worst = 0.0
    for x, y in zip(a, b):
        for k in range(4):
            scale = float(np.max(np.abs(y[k]))) or 1.0
            worst = max(worst, float(np.max(np.abs(x[k] - y[k]))) / scale)
    return worst


def interior(case, fields):
    g = case.ghost
    return [U[:, g:-g, g:-g] for U in fields]
