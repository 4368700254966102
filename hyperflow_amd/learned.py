"""Learned coarse-grid correction: data generation, training and evaluation.

The workload: MUSCL-Hancock + Van Leer + HLLC in FP32 on [0,1]^2 with
transmissive boundaries and seeded quadrant states with a sinusoidal density
perturbation (the 8192^2 benchmark's initial data at other resolutions).  A
coarse 1024^2 step is corrected by the CNN of include/hyperflow_learned.h
toward the 2x2 block average of a 2048^2 solution.

All runs use a fixed Courant number: dt = COURANT / n for an n x n grid
(dx = 1/n), COURANT = 0.4 / 6.6 -- 6.6 bounds |u| + c over every seeded
initial state (rho >= 0.09, p <= 2, |u|, |v| <= 1) -- so one coarse step spans
exactly two fine steps.  The correction is applied every K coarse steps; a
training pair is
    input  P = S_coarse^K(V)                 (K plain coarse steps)
    target T = R(U_fine(t + K dt_coarse))    (2K fine steps, restricted)
with R the conservative 2x2 block average and V the coarse state at time t:
R(U_fine(t)) in the first round (teacher forcing), the corrected coarse state
of the current network in later rounds (so the inputs follow the states the
corrected solver actually visits).

Training runs on the GPU with the library's own kernels (hf_net_train_step:
BF16 MFMA forward and backward, Adam).  torch only provides device memory.
A second stage (learned_unroll.py) trains the network through the coarse
solver itself, with the rollout's own boundaries and fallbacks.

    python -m hyperflow_amd.learned train [--seeds 40] [--every 4] [--rounds 2] [--steps 30000]
    python -m hyperflow_amd.learned evaluate [--weights hyperflow_amd/learned_weights.npz]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
# torch provides the device buffers here.  It is imported before the library
# is loaded: torch ships its own HIP runtime under the same soname, and a
# process must use a single HIP runtime, the one loaded first.
import torch

from . import autotune
from . import case as C
from . import native

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "learned_weights.npz")
COURANT = 0.4 / 6.6
COARSE, FINE = 1024, 2048
GAMMA = 1.4
SHOCK_THRESHOLD = 0.5
MAX_OUTPUT = 8.0          # fall back beyond 8 standard deviations of the training residual


def dt_for(n: int) -> float:
    return COURANT / n


def make_case(n: int, seed: int) -> C.Case:
    return C.muscl_quadrants(seed, n)


def solver(n: int, seed: int, fields=None) -> native.Solver:
    case = make_case(n, seed)
    plan = autotune.cached_params(case, "fp32", "uniform")
    s = native.Solver(case, precision="fp32", dt_mode="uniform", **plan)
    if fields is not None:
        s.upload(0, fields)
    return s


def ensure_tuned(resolutions, log=print) -> None:
    """Autotune the work plan of every resolution's class (FP32, one block)
    that has no entry of its own yet."""
    for n in resolutions:
        case = make_case(n, 0)
        if autotune.has_class(case, "fp32", "uniform"):
            continue
        r = autotune.tune(case, C.perturbed_fields(case, 0), precision="fp32", dt_mode="uniform", label=f"learned-{n}")
        log(f"tuned {n}^2: tile_rows {r['tile_rows']} tail_rows {r['tail_rows']} ({r['ms_per_step']:.4f} ms/step)")


# ------------------------------------------------------------------ data --
def remap_matrix(n_from: int, n_to: int) -> np.ndarray:
    """Conservative remap of cell averages between uniform grids on [0,1]:
    M[i, j] = |cell i of n_to  ∩  cell j of n_from| * n_to."""
    a = np.arange(n_to + 1) / n_to
    b = np.arange(n_from + 1) / n_from
    lo = np.maximum(a[:-1, None], b[None, :-1])
    hi = np.minimum(a[1:, None], b[None, 1:])
    return np.clip(hi - lo, 0.0, None) * n_to


def fine_ic(seed: int) -> np.ndarray:
    """Interior of the seeded initial state on the 2048^2 grid, [4][n][n]."""
    case = make_case(FINE, seed)
    g = case.ghost
    return C.perturbed_fields(case, seed)[0][:, g:-g, g:-g]


def ic_on(n: int, ic_fine: np.ndarray) -> np.ndarray:
    """Initial state on an n x n grid: the conservative remap of the 2048^2
    initial state (for n = 1024 exactly its 2x2 block average), padded with
    zero ghost layers (the first ghost fill sets them)."""
    if n == FINE:
        U = ic_fine
    elif n * 2 == FINE:
        U = 0.25 * ((ic_fine[:, 0::2, 0::2] + ic_fine[:, 0::2, 1::2]) + (ic_fine[:, 1::2, 0::2] + ic_fine[:, 1::2, 1::2]))
    else:
        M = remap_matrix(FINE, n)
        U = np.stack([M @ ic_fine[v] @ M.T for v in range(4)])
    g = 2
    out = np.zeros((4, n + 2 * g, n + 2 * g))
    out[:, g:-g, g:-g] = U
    return out


def make_pairs(seeds, K: int = 4, n_corr: int = 50, net=None, log=print):
    """Training pairs along coarse trajectories, device tensors [N][4][1024][1024].

    Per seed the fine run (2048^2) and a coarse run (1024^2, started from the
    restricted fine state) advance side by side; every K coarse steps the
    coarse state is an input P and the restricted fine state at the same time
    the target T.  Without a network the coarse run is then reset to the
    target (teacher forcing); with one it continues from its own corrected
    state, so the inputs follow the distribution the corrected solver visits."""
    n = len(seeds) * n_corr
    valid = np.ones(n, bool)
    P = torch.empty((n, 4, COARSE, COARSE), dtype=torch.float32, device="cuda")
    T = torch.empty_like(P)
    dtc, dtf = dt_for(COARSE), dt_for(FINE)
    assert dtf * 2 == dtc
    coarse = solver(COARSE, seeds[0])
    t0 = time.perf_counter()
    i = 0
    with ThreadPoolExecutor(max_workers=2) as pool:
        nxt = pool.submit(fine_ic, seeds[0])
        for k, seed in enumerate(seeds):
            icf = nxt.result()
            if k + 1 < len(seeds):
                nxt = pool.submit(fine_ic, seeds[k + 1])
            fine = solver(FINE, seed, ic_on(FINE, icf))
            coarse.upload(0, ic_on(COARSE, icf))
            for j in range(n_corr):
                coarse.run_fixed(K, dtc, graph=True)
                coarse.state_to_device(0, P[i])
                fine.run_fixed(2 * K, dtf, graph=True)
                fine.restrict2(0, T[i])
                # P[i] was written on the coarse solver's stream, T[i] on the
                # fine one's: both must be complete before torch's stream reads
                coarse.synchronize()
                fine.synchronize()
                if net is not None and not bool(torch.isfinite(P[i]).all()):
                    valid[i] = False          # the rollout broke down: restart it from the target
                if net is None or not valid[i]:
                    coarse.state_from_device(0, T[i])
                else:
                    coarse.correct(net)
                coarse.synchronize()
                i += 1
            fine.close()
            if (k + 1) % 10 == 0:
                log(f"pairs: {i}/{n} ({time.perf_counter() - t0:.1f} s)")
    coarse.close()
    torch.cuda.synchronize()
    if not valid.all():
        log(f"pairs: {int((~valid).sum())} rollout inputs were not finite and are left out")
    return P, T, np.flatnonzero(valid)


def norm_stats(Ps, Ts, keeps=None, sample: int = 256):
    """Per-variable input mean/scale and output scale (std of T - P) over the
    kept pairs of one or more datasets (a seeded sample of at most `sample`
    pairs)."""
    Ps = Ps if isinstance(Ps, (list, tuple)) else [Ps]
    Ts = Ts if isinstance(Ts, (list, tuple)) else [Ts]
    keeps = keeps or [np.arange(p.shape[0]) for p in Ps]
    idx = [(d, int(i)) for d, k in enumerate(keeps) for i in k]
    rng = np.random.default_rng(1)
    pick = [idx[j] for j in rng.choice(len(idx), min(sample, len(idx)), replace=False)]
    s1 = np.zeros(4)
    s2 = np.zeros(4)
    r2 = np.zeros(4)
    for d, i in pick:
        p, t = Ps[d][i].double(), Ts[d][i].double()
        s1 += p.mean(dim=(1, 2)).cpu().numpy()
        s2 += (p * p).mean(dim=(1, 2)).cpu().numpy()
        r2 += ((t - p) ** 2).mean(dim=(1, 2)).cpu().numpy()
    m = s1 / len(pick)
    std = np.sqrt(np.maximum(s2 / len(pick) - m * m, 0.0))
    out = np.sqrt(r2 / len(pick))
    return m, 1.0 / np.maximum(std, 1e-12), np.maximum(out, 1e-12)


def outliers(P, T, keep, out_scale, limit: float = 50.0):
    """Kept pairs whose largest normalised residual |T - P| / out_scale is
    below limit (rollout inputs far from the data the network was fitted on
    would dominate the loss)."""
    sc = torch.tensor(np.asarray(out_scale, np.float32), device="cuda")[:, None, None]
    good = [int(i) for i in keep if float(((T[i] - P[i]).abs() / sc).max()) < limit]
    return np.array(good, dtype=np.int64)


# -------------------------------------------------------------- training --
def train(P, T, steps: int = 20000, crop: int = 256, lr: float = 1e-3, seed: int = 0, log=print, init=None,
          keep=None):
    """Adam on random crops of random pairs; cosine learning-rate decay.  P and
    T are tensors [N][4][ny][nx] or lists of them (datasets of several rounds,
    sampled pair-uniformly); `keep` lists the usable pair indices of each.
    `init` (a Net) gives the starting parameters and normalisation.  Returns
    the network and the loss log."""
    Ps = P if isinstance(P, (list, tuple)) else [P]
    Ts = T if isinstance(T, (list, tuple)) else [T]
    if keep is None:
        keep = [np.arange(p.shape[0]) for p in Ps]
    elif not isinstance(keep, (list, tuple)):
        keep = [keep]
    sizes = np.array([len(k) for k in keep])
    if init is None:
        mean, iscale, oscale = norm_stats(Ps, Ts, list(keep))
    else:
        d = init.desc
        mean, iscale, oscale = d["in_mean"], d["in_scale"], d["out_scale"]
    net = native.Net(in_mean=mean, in_scale=iscale, out_scale=oscale, gamma=GAMMA, shock_threshold=SHOCK_THRESHOLD,
                     seed=seed, max_output=MAX_OUTPUT)
    if init is not None:
        net.set_params(init.get_params())
    rng = np.random.default_rng(seed)
    ny, nx = Ps[0].shape[2], Ps[0].shape[3]
    loss_dev = torch.zeros(steps, dtype=torch.float32, device="cuda")
    up = torch.empty((4, crop, crop), dtype=torch.float32, device="cuda")
    tp = torch.empty_like(up)
    t0 = time.perf_counter()
    for it in range(steps):
        i = int(rng.integers(sizes.sum()))
        dset = int(np.searchsorted(np.cumsum(sizes), i, side="right"))
        i = int(keep[dset][i - int(sizes[:dset].sum())])
        y0 = int(rng.integers(ny - crop + 1))
        x0 = int(rng.integers(nx // 64 - crop // 64 + 1)) * 64
        up.copy_(Ps[dset][i, :, y0:y0 + crop, x0:x0 + crop])
        tp.copy_(Ts[dset][i, :, y0:y0 + crop, x0:x0 + crop])
        a = lr * 0.5 * (1.0 + math.cos(math.pi * it / steps))
        net.train_step(up, tp, a, loss_dev[it:it + 1])
        if (it + 1) % 5000 == 0:
            torch.cuda.synchronize()
            recent = float(loss_dev[it - 4999:it + 1].mean())
            log(f"train: step {it + 1}/{steps} loss {recent:.4e} ({time.perf_counter() - t0:.1f} s)")
    torch.cuda.synchronize()
    return net, loss_dev.cpu().numpy()


def full_loss(net, P, T, idx):
    """Mean loss over whole fields idx (network output vs the normalised residual)."""
    vals = [net.gradients(P[i], T[i])[0] for i in idx]
    return float(np.mean(vals))


def save(net, path: str, meta: dict) -> None:
    d = net.desc
    np.savez(path, params=net.get_params(), in_mean=np.array(d["in_mean"], np.float32),
             in_scale=np.array(d["in_scale"], np.float32), out_scale=np.array(d["out_scale"], np.float32),
             gamma=np.float32(d["gamma"]), shock_threshold=np.float32(d["shock_threshold"]),
             max_output=np.float32(d["max_output"]), meta=np.array(json.dumps(meta)))


def load(path: str = WEIGHTS, shock_threshold=None, max_output=None) -> native.Net:
    """The network saved at `path`; shock_threshold / max_output override the
    stored fallback settings."""
    z = np.load(path, allow_pickle=False)
    th = float(z["shock_threshold"]) if shock_threshold is None else shock_threshold
    mo = (float(z["max_output"]) if "max_output" in z.files else 0.0) if max_output is None else max_output
    net = native.Net(in_mean=z["in_mean"], in_scale=z["in_scale"], out_scale=z["out_scale"],
                     gamma=float(z["gamma"]), shock_threshold=th, max_output=mo)
    net.set_params(z["params"])
    return net


def load_meta(path: str = WEIGHTS) -> dict:
    z = np.load(path, allow_pickle=False)
    return json.loads(str(z["meta"]))


# ------------------------------------------------------------ evaluation --
def density_l1(rho: np.ndarray, truth: np.ndarray) -> float:
    """L1 norm of the density error on the 1024^2 truth grid (rho remapped)."""
    n = rho.shape[0]
    if n != truth.shape[0]:
        M = remap_matrix(n, truth.shape[0])
        rho = M @ rho.astype(np.float64) @ M.T
    return float(np.mean(np.abs(rho.astype(np.float64) - truth.astype(np.float64))))


def _timed(fn):
    torch.cuda.synchronize()
    t = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t


def run_plain(n: int, icf: np.ndarray, coarse_steps: int, seed: int = 0):
    """Plain run to t = coarse_steps * dt_coarse on an n x n grid from the
    remapped initial state; returns (final density [n][n], seconds), timed as
    the second of two identical runs (the first captures the HIP graph)."""
    steps = coarse_steps * n // COARSE
    assert steps * COARSE == coarse_steps * n
    ic = ic_on(n, icf)
    s = solver(n, seed)
    secs = []
    for _ in range(2):
        s.upload(0, ic)
        s.synchronize()
        secs.append(_timed(lambda: (s.run_fixed(steps, dt_for(n), graph=True), s.synchronize())))
    out = torch.empty((4, n, n), dtype=torch.float32, device="cuda")
    s.state_to_device(0, out)
    s.synchronize()
    rho = out[0].cpu().numpy()
    s.close()
    return rho, secs[-1]


def run_learned(net, icf: np.ndarray, coarse_steps: int, every: int, seed: int = 0, start: int = 0):
    """Coarse run with the correction after every `every` steps from step
    `start` on; returns (density, seconds, fallback fraction over all
    corrections)."""
    ic = ic_on(COARSE, icf)
    s = solver(COARSE, seed)
    dtc = dt_for(COARSE)

    def go(count=False):
        fb = 0
        if start:
            s.run_fixed(start, dtc, graph=True)
        for k in range(start, coarse_steps, every):
            s.run_fixed(min(every, coarse_steps - k), dtc, graph=True)
            fb += s.correct(net, count=count) or 0
        s.synchronize()
        return fb

    secs = []
    for _ in range(2):
        s.upload(0, ic)
        s.synchronize()
        secs.append(_timed(go))
    s.upload(0, ic)
    fb = go(count=True)          # same run again, counting fallbacks (not timed)
    out = torch.empty((4, COARSE, COARSE), dtype=torch.float32, device="cuda")
    s.state_to_device(0, out)
    s.synchronize()
    rho = out[0].cpu().numpy()
    s.close()
    ncorr = max(1, (coarse_steps - start + every - 1) // every)
    return rho, secs[-1], fb / float(ncorr * COARSE * COARSE)


def truth(icf: np.ndarray, coarse_steps: int, seed: int = 0):
    """2x2 block average of the plain 2048^2 solution (density) and its time."""
    s = solver(FINE, seed)
    ic = ic_on(FINE, icf)
    secs = []
    for _ in range(2):
        s.upload(0, ic)
        s.synchronize()
        secs.append(_timed(lambda: (s.run_fixed(2 * coarse_steps, dt_for(FINE), graph=True), s.synchronize())))
    out = torch.empty((4, COARSE, COARSE), dtype=torch.float32, device="cuda")
    s.restrict2(0, out)
    s.synchronize()
    rho = out[0].cpu().numpy()
    s.close()
    return rho, secs[-1]


def evaluate(net, seeds, coarse_steps: int = 200, resolutions=(1024, 1152, 1280, 1536, 1792), every=(8,),
             start: int = 0, log=print):
    """Density L1 against the restricted fine solution, and wall times.

    Every run starts from the conservative remap of the same 2048^2 initial
    state.  The plain resolution n* that matches the learned L1, and its wall
    time, are interpolated in log-log between the two measured plain runs that
    bracket the learned L1 (outside the measured range: from least-squares
    fits, L1 ~ n^b and time ~ n^3, n^2 cells times n steps at a fixed Courant
    number); the learned speed-up at equal L1 is time(n*) / time(learned).
    The learned run corrects every `every` coarse steps from step `start` on.
    Resolutions must be multiples of 128 (whole numbers of steps)."""
    rows = []
    for seed in seeds:
        icf = fine_ic(seed)
        tr, t_fine = truth(icf, coarse_steps, seed)
        r = {"seed": seed, "fine_seconds": t_fine, "plain": {}, "learned": {}}
        for n in resolutions:
            rho, t = run_plain(n, icf, coarse_steps, seed)
            r["plain"][n] = {"l1": density_l1(rho, tr), "seconds": t}
        for e in every:
            rho, t, fb = run_learned(net, icf, coarse_steps, e, seed, start)
            r["learned"][e] = {"l1": density_l1(rho, tr), "seconds": t, "fallback_fraction": fb}
        log(f"eval seed {seed}: " + json.dumps(r, default=float))
        rows.append(r)
    summary = summarize(rows, resolutions, every)
    return rows, summary


def equal_l1_point(ns, l1s, ts, l1):
    """(n*, time(n*)) of the plain runs at density L1 `l1`, interpolated
    linearly in log-log between the two measured resolutions whose L1 brackets
    it; (nan, nan) when no neighbouring pair brackets it."""
    ns, l1s, ts = (np.asarray(v, np.float64) for v in (ns, l1s, ts))
    for i in range(len(ns) - 1):
        hi, lo = l1s[i], l1s[i + 1]
        if hi >= l1 >= lo and hi > lo:
            w = np.log(hi / l1) / np.log(hi / lo)
            n = np.exp(np.log(ns[i]) + w * (np.log(ns[i + 1]) - np.log(ns[i])))
            t = np.exp(np.log(ts[i]) + w * (np.log(ts[i + 1]) - np.log(ts[i])))
            return float(n), float(t)
    return float("nan"), float("nan")


def summarize(rows, resolutions, every):
    ns = np.array(resolutions, np.float64)
    l1p = np.array([[r["plain"][n]["l1"] for n in resolutions] for r in rows]).mean(axis=0)
    tp = np.array([[r["plain"][n]["seconds"] for n in resolutions] for r in rows]).mean(axis=0)
    out = {"resolutions": list(resolutions), "plain_l1": l1p.tolist(), "plain_seconds": tp.tolist(),
           "fine_seconds": float(np.mean([r["fine_seconds"] for r in rows])), "learned": {}}
    # fit log L1 = a + b log n and log t = c + 3 log n (least squares on c)
    b, a = np.polyfit(np.log(ns), np.log(l1p), 1)
    c = float(np.mean(np.log(tp) - 3.0 * np.log(ns)))
    for e in every:
        l1 = float(np.mean([r["learned"][e]["l1"] for r in rows]))
        t = float(np.mean([r["learned"][e]["seconds"] for r in rows]))
        n_fit = float(np.exp((np.log(l1) - a) / b)) if b < 0 else float("nan")
        t_fit = float(np.exp(c + 3.0 * np.log(n_fit))) if np.isfinite(n_fit) else float("nan")
        n_eq, t_eq = equal_l1_point(ns, l1p, tp, l1)
        if not np.isfinite(n_eq):
            n_eq, t_eq = n_fit, t_fit
        out["learned"][e] = {"l1": l1, "seconds": t, "plain_n_at_equal_l1": n_eq,
                             "plain_seconds_at_equal_l1": t_eq, "speedup_at_equal_l1": t_eq / t,
                             "plain_n_at_equal_l1_fit": n_fit, "speedup_at_equal_l1_fit": t_fit / t,
                             "l1_gain_vs_plain_1024": float(l1p[0] / l1), "speedup_vs_plain_2048": out["fine_seconds"] / t}
    out["l1_slope"] = float(b)
    return out


def bench_net(n: int = 1024, reps: int = 50) -> dict:
    """Average time of the forward pass and of a training step on an n x n
    state (random weights, synthetic smooth state)."""
    net = native.Net(in_mean=[1, 0, 0, 2.5], in_scale=[1, 1, 1, 1], out_scale=[0.01] * 4, seed=1)
    yy, xx = torch.meshgrid(torch.linspace(0, 1, n, device="cuda"), torch.linspace(0, 1, n, device="cuda"),
                            indexing="ij")
    u = torch.stack([1 + 0.2 * torch.sin(6 * xx), 0.1 * torch.cos(4 * yy), 0.1 * torch.sin(5 * xx),
                     2.5 + 0.1 * torch.cos(3 * (xx + yy))]).contiguous()
    t = (u * 1.001).contiguous()
    y = torch.zeros_like(u)
    for _ in range(3):
        net.forward(u, y)
        net.train_step(u, t, 1e-6)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        net.forward(u, y)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(reps):
        net.train_step(u, t, 1e-6)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    flops = 2 * 9 * sum(ci * co for ci, co in ((4, 32), (32, 32), (32, 32), (32, 32), (32, 4))) * n * n
    return {"n": n, "forward_ms": (t1 - t0) / reps * 1e3, "train_step_ms": (t2 - t1) / reps * 1e3,
            "forward_tflops": flops / ((t1 - t0) / reps) / 1e12}


# ------------------------------------------------------------------ main --
def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m hyperflow_amd.learned")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--seeds", type=int, default=40, help="training seeds 0..n-1 (each a 2048^2 run)")
    t.add_argument("--every", type=int, default=4, help="coarse steps K of a training pair")
    t.add_argument("--corrections", type=int, default=50, help="pairs per seed and round")
    t.add_argument("--rounds", type=int, default=1, help="1 = teacher forcing only; more add rollout rounds")
    t.add_argument("--apply-every", type=int, default=8, help="correction interval stored for evaluation")
    t.add_argument("--apply-start", type=int, default=24, help="first corrected step stored for evaluation")
    t.add_argument("--steps", type=int, default=30000)
    t.add_argument("--crop", type=int, default=256)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--lr-later", type=float, default=3e-4, help="learning rate of the rollout rounds")
    t.add_argument("--holdout-seeds", type=int, default=2)
    t.add_argument("--out", default=WEIGHTS)
    t.add_argument("--save-rounds", default=None, help="directory for the weights after every round")
    b = sub.add_parser("bench", help="time the network's forward pass and training step")
    b.add_argument("--n", type=int, default=1024)
    b.add_argument("--reps", type=int, default=50)
    e = sub.add_parser("evaluate")
    e.add_argument("--weights", default=WEIGHTS)
    e.add_argument("--seeds", type=int, nargs="+", default=[5000, 5001, 5002])
    e.add_argument("--coarse-steps", type=int, default=200)
    e.add_argument("--every", type=int, nargs="+", default=None, help="default: the interval stored with the weights")
    e.add_argument("--start", type=int, default=None, help="first corrected step; default: stored with the weights")
    e.add_argument("--shock-threshold", type=float, default=None)
    e.add_argument("--max-output", type=float, default=None, help="fallback bound on |network output|")
    e.add_argument("--resolutions", type=int, nargs="+", default=[1024, 1152, 1280, 1536, 1792])
    e.add_argument("--json", default=None)
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    if args.cmd == "bench":
        log(json.dumps(bench_net(args.n, args.reps)))
        return 0
    if args.cmd == "train":
        ensure_tuned((COARSE, FINE), log=log)
        seeds = list(range(args.seeds))
        train_seeds, hold_seeds = seeds[:len(seeds) - args.holdout_seeds], seeds[len(seeds) - args.holdout_seeds:]
        K, nc = args.every, args.corrections
        Ps, Ts, keeps, net, meta_rounds = [], [], [], None, []
        hold = make_pairs(hold_seeds, K, nc, None, log=log)
        for rnd in range(args.rounds):
            P, T, keep = make_pairs(train_seeds, K, nc, net, log=log)
            if net is not None:
                keep = outliers(P, T, keep, norm_stats(Ps, Ts, keeps)[2] * 4.0)
                log(f"round {rnd}: {len(keep)} of {P.shape[0]} rollout pairs kept")
            Ps.append(P)
            Ts.append(T)
            keeps.append(keep)
            net, losses = train(Ps, Ts, steps=args.steps, crop=args.crop, lr=args.lr if rnd == 0 else args.lr_later,
                                log=log, keep=keeps)
            meta_rounds.append({"round": rnd, "pairs": int(sum(p.shape[0] for p in Ps)),
                                "loss_first_2000": float(losses[:2000].mean()),
                                "loss_last_2000": float(losses[-2000:].mean()),
                                "holdout_loss": full_loss(net, hold[0], hold[1], range(0, hold[0].shape[0], 5))})
            log("train: " + json.dumps(meta_rounds[-1]))
            if args.save_rounds:
                save(net, os.path.join(args.save_rounds, f"learned_round{rnd}.npz"),
                     {"every": K, "round": rnd, **meta_rounds[-1]})
        zero = native.Net(**{k: net.desc[k] for k in ("in_mean", "in_scale", "out_scale")})
        zero.set_params(np.zeros(native.Net.param_count(), np.float32))
        meta = {"train_seeds": [train_seeds[0], train_seeds[-1]], "holdout_seeds": hold_seeds, "trained_every": K,
                "every": args.apply_every, "start": args.apply_start,
                "corrections_per_seed": nc, "rounds": meta_rounds, "steps_per_round": args.steps, "crop": args.crop,
                "lr": args.lr, "courant": COURANT, "coarse": COARSE, "fine": FINE,
                "holdout_loss_no_correction": full_loss(zero, hold[0], hold[1], range(0, hold[0].shape[0], 5))}
        log("train: " + json.dumps(meta))
        save(net, args.out, meta)
        return 0
    ensure_tuned(sorted(set((COARSE, FINE) + tuple(args.resolutions))), log=log)
    net = load(args.weights, args.shock_threshold, args.max_output)
    meta = load_meta(args.weights)
    every = args.every or [int(meta.get("every", 1))]
    start = args.start if args.start is not None else int(meta.get("start", 0))
    rows, summary = evaluate(net, args.seeds, args.coarse_steps, resolutions=tuple(args.resolutions),
                             every=tuple(every), start=start, log=log)
    summary["start"] = start
    log("summary: " + json.dumps(summary))
    if args.json:
        with open(args.json, "w") as f:
            json.dump({"rows": rows, "summary": summary}, f, indent=1, default=float)
    return 0


if __name__ == "__main__":
    sys.exit(main())
