"""Training the learned correction through the coarse solver (solver in the loop).

The pairs of learned.make_pairs teach the network to undo K coarse steps of
error from a state that lies exactly on the restricted fine trajectory; a
rollout then feeds it the states its own earlier corrections produced, and
the single-step gain shrinks over 25 corrections.  Here the loss runs
through J corrections of K coarse steps each:

    V_0 = R(U_fine(t)),   V_j = C(S^K(V_{j-1})),
    loss = sum_j mean|V_j - R(U_fine(t + j K dt))| / out_scale

with S the FP32 MUSCL-Hancock + Van Leer + HLLC step at the fixed coarse dt
and C the corrected state (network output, then the fallbacks of
hf_correct: non-positive density or pressure, the shock sensor, the output
bound).  Gradients reach the parameters through every step and every
earlier correction.

S and C are written here as differentiable torch expressions.  This module is
a training tool only: the solver and the corrections that runs apply are the
library's kernels (hf_run_fixed, hf_correct), and the trained parameters go
back into the library's network (hf_net_set_params).  The step restates the
FP32 formulation of the library's kernel (hf_physics.h, namespace f32) with
the same operations, unfused; in float64 it agrees with the reference's
MUSCL-Hancock residual (src/fvm_2d_cart/scheme/muscl_hancock/
muscl_hancock.cpp:243-376) to rounding.  Its Van Leer and MinBee limiters
follow vanleer.cpp:29-36 and minbee.cpp:29-38 with the ratio's tolerance
clamp of muscl_hancock.cpp:78-94, its Riemann solver riemann_hllc.cpp:25-212.

Samples.  With --whole (how the shipped network was trained) a sample is a
whole 1024^2 domain: the steps take the solver's transmissive ghosts and the
network the zero padding hf_correct gives it at the domain edges, so training
sees every cell as runs do.  Otherwise a sample is a square crop: a step needs
two cells of neighbours on every side and each 3x3 convolution one, so the
state shrinks by 2 cells per side per step and 5 per correction ("valid"
evaluation) and the loss sees the cells whose domain of dependence lies in
the crop; crops never train the cells along the domain edges, and what they
learned did not carry over to runs.

    python -m hyperflow_amd.learned_unroll train --init hyperflow_amd/learned_weights.npz --out w.npz \
        --whole --every 16 --unroll 1 --lr 3e-4 --seconds 420
"""
from __future__ import annotations

import argparse
import math
import os
import sys
import time

# MIOpen's exhaustive kernel search for every new convolution shape (three
# per correction here) takes minutes on the first backward pass; the
# immediate-mode choice takes under a second and trains as fast afterwards.
os.environ.setdefault("MIOPEN_FIND_MODE", "FAST")

import numpy as np
import torch
import torch.nn.functional as F
from torch.utils.checkpoint import checkpoint

LAYERS = [(4, 32), (32, 32), (32, 32), (32, 32), (32, 4)]
TOL = 1e-6            # the limiter ratio's clamp (muscl_hancock.cpp:8-11)
STEP_HALO = 2         # cells consumed per side by one step
NET_HALO = 5          # cells consumed per side by the five 3x3 convolutions


# ------------------------------------------------------------------ step --
def _prims(U0, U1, U2, U3, g1):
    ri = 1.0 / U0
    u = U1 * ri
    v = U2 * ri
    p = g1 * (U3 - 0.5 * (U1 * u + U2 * v))
    return ri, u, v, p


def _flux(X, U0, U1, U2, U3, u, v, p):
    if X:
        return (U1, U1 * u + p, U1 * v, u * (U3 + p))
    return (U2, U2 * u, U2 * v + p, v * (U3 + p))


def _clamp_tol(d):
    return torch.where(d.abs() <= TOL, torch.where(d >= 0, TOL, -TOL), d)


def _slope(um, uc, up, vanleer):
    """xi(r) (dl + dr) / 2 with r from the clamped deltas; zero unless r > 0."""
    dl, dr = uc - um, up - uc
    dlc, drc = _clamp_tol(dl), _clamp_tol(dr)
    same = dlc * drc > 0
    le1 = dlc.abs() <= drc.abs()
    if vanleer:      # 2 min(r, 1) / (1 + r)
        num = torch.where(le1, dlc, drc)
        den = dlc + drc
    else:            # MinBee: r for r <= 1, 2 / (1 + r) above
        num = torch.where(le1, 0.5 * dlc, drc)
        den = torch.where(le1, drc, drc + dlc)
    den = torch.where(same, den, torch.ones_like(den))   # keeps the unused branch finite for autograd
    return torch.where(same, (dl + dr) * num / den, torch.zeros_like(dl))


def _reconstruct(U, axis, h, g1, vanleer):
    """Face states of the cells 1..n-2 along `axis` of U[..., 4, H, W] with
    the half-step predictor of that direction: lo/hi = U -+ s/2, then
    += h (F(lo) - F(hi)).  Returns the two as tensors like U."""
    n = U.shape[axis] - 2
    um, uc, up = (U.narrow(axis, o, n) for o in range(3))
    s = _slope(um, uc, up, vanleer)
    lo, hi = uc - 0.5 * s, uc + 0.5 * s
    X = axis == U.dim() - 1
    ql, qh = lo.unbind(-3), hi.unbind(-3)
    _, ul, vl, pl = _prims(*ql, g1)
    _, ur, vr, pr = _prims(*qh, g1)
    d = h * (torch.stack(_flux(X, *ql, ul, vl, pl), -3) - torch.stack(_flux(X, *qh, ur, vr, pr), -3))
    return lo + d, hi + d


def _hllc(X, L, R, gamma, g1):
    """HLLC flux (Davis speeds) between face states L and R, tensors
    [..., 4, h, w]; returns the flux as one tensor of the same shape.  As in
    the kernel, a NaN wave speed gives the zero flux."""
    riL, uL, vL, pL = _prims(*L.unbind(-3), g1)
    riR, uR, vR, pR = _prims(*R.unbind(-3), g1)
    a2L, a2R = gamma * pL * riL, gamma * pR * riR
    # a negative (or NaN) a^2 is a NaN speed, which gives the zero flux; the
    # square root only sees valid arguments, so no NaN reaches the gradient
    ok = (a2L >= 0) & (a2R >= 0)
    aL = torch.sqrt(torch.where(ok, a2L, torch.ones_like(a2L)))
    aR = torch.sqrt(torch.where(ok, a2R, torch.ones_like(a2R)))
    qL, qR = (uL, uR) if X else (vL, vR)
    S_L = torch.minimum(qL - aL, qR - aR)
    S_R = torch.maximum(qL + aL, qR + aR)
    mL, mR = L[..., 0, :, :] * (S_L - qL), R[..., 0, :, :] * (S_R - qR)
    den = mL - mR
    ok = ok & (den != 0)
    Ss = (mL * qL - mR * qR + (pR - pL)) / torch.where(ok, den, torch.ones_like(den))
    ok = ok & torch.isfinite(S_L + S_R + Ss)
    useR = (S_L < 0) & (Ss < 0)
    zero = torch.zeros_like(S_L)
    SK = torch.where(useR, S_R, S_L)
    w = torch.where(useR, torch.where(S_R > 0, S_R, zero), torch.where(S_L < 0, S_L, zero))
    K = torch.where(useR.unsqueeze(-3), R, L)
    uK, vK, pK, mK = (torch.where(useR, b, a) for a, b in ((uL, uR), (vL, vR), (pL, pR), (mL, mR)))
    dK, U1K, U2K, EK = K.unbind(-3)
    qK, tK = (uK, vK) if X else (vK, uK)
    FK = torch.stack(_flux(X, dK, U1K, U2K, EK, uK, vK, pK), -3)
    star = (w != 0) & (SK != Ss)
    rS = 1.0 / torch.where(star, SK - Ss, torch.ones_like(SK))
    us0 = mK * rS
    us3 = rS * ((SK - qK) * EK + (Ss - qK) * (mK * Ss + pK))
    # U*_K with the normal and tangential momenta in their slots
    un, ut = us0 * Ss, us0 * tK
    Us = torch.stack((us0, un if X else ut, ut if X else un, us3), -3)
    F = torch.where(star.unsqueeze(-3), FK + w.unsqueeze(-3) * (Us - K), FK)
    return torch.where(ok.unsqueeze(-3), F, torch.zeros_like(F))


def muscl_step(U, dt, dx, gamma=1.4, dt_pred=None, limiter="vanleer", dy=None):
    """One MUSCL-Hancock step of U[..., 4, H, W] (dy defaults to dx), valid
    evaluation: returns the updated cells 2..H-3 x 2..W-3.  dt_pred, the
    predictor's dt, defaults to dt."""
    vanleer = {"vanleer": True, "minbee": False}[limiter]
    dt_pred = dt if dt_pred is None else dt_pred
    dy = dx if dy is None else dy
    g1 = gamma - 1.0
    hx, hy = (0.5 * dt_pred) / dx, (0.5 * dt_pred) / dy
    rdx, rdy = 1.0 / dx, 1.0 / dy
    H, W = U.shape[-2], U.shape[-1]
    ya, xa = U.dim() - 2, U.dim() - 1
    # x: rows of the updated cells, every column; y: columns of the updated cells
    Lx, Rx = _reconstruct(U.narrow(ya, 2, H - 4), xa, hx, g1, vanleer)
    Ly, Ry = _reconstruct(U.narrow(xa, 2, W - 4), ya, hy, g1, vanleer)
    nfx, nfy = W - 3, H - 3
    Fx = _hllc(True, Rx.narrow(xa, 0, nfx), Lx.narrow(xa, 1, nfx), gamma, g1)
    Fy = _hllc(False, Ry.narrow(ya, 0, nfy), Ly.narrow(ya, 1, nfy), gamma, g1)
    L = rdx * (Fx.narrow(xa, 0, W - 4) - Fx.narrow(xa, 1, W - 4)) + rdy * (Fy.narrow(ya, 0, H - 4) - Fy.narrow(ya, 1, H - 4))
    return U.narrow(ya, 2, H - 4).narrow(xa, 2, W - 4) + dt * L


# --------------------------------------------------------------- network --
class TorchNet(torch.nn.Module):
    """The library's network (include/hyperflow_learned.h) as torch modules:
    x = (u - in_mean) * in_scale, four 3x3 convolutions with ReLU, a last one;
    valid evaluation (the output is 5 cells smaller on every side)."""

    def __init__(self, in_mean, in_scale, out_scale):
        super().__init__()
        self.convs = torch.nn.ModuleList(torch.nn.Conv2d(ci, co, 3) for ci, co in LAYERS)
        for name, val in (("in_mean", in_mean), ("in_scale", in_scale), ("out_scale", out_scale)):
            self.register_buffer(name, torch.tensor(np.asarray(val, np.float64)).view(4, 1, 1))

    def set_params(self, params) -> None:
        """From the library's flat order: per layer W[9][ci][co] (tap =
        3 (dy+1) + (dx+1), input pixel (y+dy, x+dx)), then b[co]."""
        params = np.asarray(params, np.float64)
        off = 0
        with torch.no_grad():
            for conv, (ci, co) in zip(self.convs, LAYERS):
                Wn = params[off:off + 9 * ci * co].reshape(3, 3, ci, co)
                off += 9 * ci * co
                conv.weight.copy_(torch.from_numpy(Wn.transpose(3, 2, 0, 1).copy()))
                conv.bias.copy_(torch.from_numpy(params[off:off + co].copy()))
                off += co
        assert off == params.size

    def get_params(self) -> np.ndarray:
        out = []
        for conv in self.convs:
            out.append(conv.weight.detach().double().cpu().numpy().transpose(2, 3, 1, 0).reshape(-1))
            out.append(conv.bias.detach().double().cpu().numpy())
        return np.concatenate(out).astype(np.float32)

    def forward(self, u, same: bool = False):
        """Valid evaluation, or with same=True zero padding around every
        layer's input (the library's convention at the domain edges)."""
        a = (u - self.in_mean.to(u.dtype)) * self.in_scale.to(u.dtype)
        for l, conv in enumerate(self.convs):
            a = F.conv2d(a, conv.weight, conv.bias, padding=1 if same else 0)
            if l < len(self.convs) - 1:
                a = F.relu(a)
        return a


def _pressure(U, g1):
    return g1 * (U[..., 3, :, :] - 0.5 * (U[..., 1, :, :] ** 2 + U[..., 2, :, :] ** 2) / U[..., 0, :, :])


def correct(net, U, gamma=1.4, shock_threshold=0.5, max_output=8.0, whole: bool = False):
    """Corrected state of U[..., 4, H, W] on its cells 5..H-6 x 5..W-6 (with
    whole=True: on every cell, U being the whole domain, as hf_correct does
    it: zero padding in the network, the sensor's neighbours clamped at the
    edges): the network's correction, except where the corrected density or
    pressure is not positive and finite, the shock sensor fires on the plain
    state, or some |output| exceeds max_output (the rules of hf_correct)."""
    g1 = gamma - 1.0
    y = net(U, same=whole)
    H, W = U.shape[-2], U.shape[-1]
    h = 0 if whole else NET_HALO
    Uc = U[..., h:H - h, h:W - h]
    n = Uc + net.out_scale.to(U.dtype) * y
    with torch.no_grad():
        p = _pressure(n, g1)
        bad = ~(n[..., 0, :, :] > 0) | ~(p > 0) | ~torch.isfinite(p) | ~torch.isfinite(n).all(dim=-3)
        if shock_threshold > 0:
            if whole:
                P = F.pad(_pressure(U, g1).unsqueeze(-3), (1, 1, 1, 1), mode="replicate").squeeze(-3)
            else:
                P = _pressure(U[..., h - 1:H - h + 1, h - 1:W - h + 1], g1)
            c = P[..., 1:-1, 1:-1]
            sx = (P[..., 1:-1, 2:] - 2 * c + P[..., 1:-1, :-2]).abs() / (P[..., 1:-1, 2:] + 2 * c + P[..., 1:-1, :-2])
            sy = (P[..., 2:, 1:-1] - 2 * c + P[..., :-2, 1:-1]).abs() / (P[..., 2:, 1:-1] + 2 * c + P[..., :-2, 1:-1])
            bad |= (sx > shock_threshold) | (sy > shock_threshold) | torch.isnan(sx) | torch.isnan(sy)
        if max_output > 0:
            bad |= (y.abs() > max_output).any(dim=-3)
    return torch.where(bad.unsqueeze(-3), Uc, n), bad


# ------------------------------------------------------------------ data --
def make_sequences(seeds, K: int, count: int, log=print):
    """Restricted fine states R(U_fine(j K dt_coarse)), j = 0..count, per seed:
    a device tensor [len(seeds)][count+1][4][1024][1024] from the library's
    2048^2 FP32 runs (HIP graphs)."""
    from . import learned as Lm
    n = Lm.COARSE
    T = torch.empty((len(seeds), count + 1, 4, n, n), dtype=torch.float32, device="cuda")
    dtf = Lm.dt_for(Lm.FINE)
    t0 = time.perf_counter()
    for k, seed in enumerate(seeds):
        icf = Lm.fine_ic(seed)
        fine = Lm.solver(Lm.FINE, seed, Lm.ic_on(Lm.FINE, icf))
        for j in range(count + 1):
            if j:
                fine.run_fixed(2 * K, dtf, graph=True)
            fine.restrict2(0, T[k, j])
        fine.synchronize()
        fine.close()
        if (k + 1) % 10 == 0:
            log(f"sequences: {k + 1}/{len(seeds)} seeds ({time.perf_counter() - t0:.1f} s)")
    torch.cuda.synchronize()
    return T


def transmissive_step(U, dt, dx, gamma=1.4):
    """One step of the whole domain U[..., 4, H, W] with transmissive
    boundaries, as the library's solver takes it: ghost layer l (l = 0 next
    to the interior) holds interior cell l, mirrored across the edge
    (transmissive_boundary.cpp:25-104).  The corner ghosts are not read."""
    ya, xa = U.dim() - 2, U.dim() - 1
    U = torch.cat([U.narrow(ya, 0, 2).flip(ya), U, U.narrow(ya, U.shape[ya] - 2, 2).flip(ya)], dim=ya)
    U = torch.cat([U.narrow(xa, 0, 2).flip(xa), U, U.narrow(xa, U.shape[xa] - 2, 2).flip(xa)], dim=xa)
    return muscl_step(U, dt, dx, gamma)


def rollout_loss(net, T, picks, K: int, J: int, crop: int, dt: float, dx: float, gamma: float,
                 shock_threshold: float, max_output: float, norm: str = "l2", whole: bool = False):
    """Mean over the picks (seed, first index, y0, x0) and the J corrections of
    the error of the corrected crop against the restricted fine state,
    normalised by out_scale: squared ("l2") or absolute ("l1").  The squared
    error is dominated by the few cells at misplaced shocks, where the
    correction falls back or cannot reach; the absolute error weights the
    smooth flow the correction can improve, and is the evaluation's norm.
    With whole=True every pick is the whole domain (y0 = x0 = 0, crop = its
    side): transmissive steps and the edge rules of hf_correct, nothing
    shrinks."""
    U = torch.stack([T[s, j, :, y0:y0 + crop, x0:x0 + crop] for s, j, y0, x0 in picks])
    step = transmissive_step if whole else muscl_step
    oscale = net.out_scale.to(U.dtype)
    total = 0.0
    off = 0
    for j in range(1, J + 1):
        for _ in range(K):
            if torch.is_grad_enabled():
                # keeps only the step's input; the backward pass recomputes
                # the step (about 2 KB of intermediates per cell otherwise)
                U = checkpoint(step, U, dt, dx, gamma, use_reentrant=False)
            else:
                U = step(U, dt, dx, gamma)
        U, _ = correct(net, U, gamma, shock_threshold, max_output, whole)
        if not whole:
            off += K * STEP_HALO + NET_HALO
        m = crop - 2 * off
        tgt = torch.stack([T[s, jj + j, :, y0 + off:y0 + off + m, x0 + off:x0 + off + m]
                           for s, jj, y0, x0 in picks]).to(U.dtype)
        e = (U - tgt) / oscale
        total = total + ((e * e).mean() if norm == "l2" else e.abs().mean())
    return total / J


# -------------------------------------------------------------- training --
def train(T, init: str, K: int = 8, J: int = 3, crop_out: int = 128, batch: int = 8, steps: int = 2000,
          lr: float = 1e-4, seed: int = 0, first: int = 0, log=print, hold=None, log_every: int = 100,
          seconds: float = 0.0, norm: str = "l1", whole: bool = False):
    """Adam on the rollout loss from the parameters and normalisation stored
    at `init`; cosine learning-rate decay, gradient norm clipped at 1.
    T: make_sequences output of the training seeds (the network lives on its
    device); hold: that of held-out seeds, whose rollout loss on fixed picks
    is logged.  With seconds > 0 the step count is cut, after 10 steps, to
    what the measured rate fits in that time.  whole=True trains on whole
    domains (the boundaries as the library's solver and hf_correct treat
    them) instead of crops.  Returns the network and a log of (step, mean
    training loss, held-out loss)."""
    from . import learned as Lm
    z = np.load(init, allow_pickle=False)
    th, mo = float(z["shock_threshold"]), float(z["max_output"])
    gamma = float(z["gamma"])
    net = TorchNet(z["in_mean"], z["in_scale"], z["out_scale"]).to(T.device)
    net.set_params(z["params"])
    opt = torch.optim.Adam(net.parameters(), lr=lr)
    crop = T.shape[-1] if whole else crop_out + 2 * J * (K * STEP_HALO + NET_HALO)
    n, count = T.shape[-1], T.shape[1] - 1
    dt, dx = Lm.dt_for(Lm.COARSE), 1.0 / Lm.COARSE
    assert count - J >= first, "sequences too short for the unroll"
    rng = np.random.default_rng(seed)

    def pick(r, nseeds):
        return [(int(r.integers(nseeds)), int(r.integers(first, count - J + 1)), int(r.integers(n - crop + 1)),
                 int(r.integers(n - crop + 1))) for _ in range(batch)]

    hold_picks = [pick(np.random.default_rng(1000 + i), hold.shape[0]) for i in range(4)] if hold is not None else []

    def hold_loss():
        if hold is None:
            return float("nan")
        with torch.no_grad():
            return float(np.mean([float(rollout_loss(net, hold, p, K, J, crop, dt, dx, gamma, th, mo, norm, whole))
                                  for p in hold_picks]))

    history = [(0, float("nan"), hold_loss())]
    if hold is not None:              # the same rollouts without corrections, for scale
        saved = [p.detach().clone() for p in net.parameters()]
        with torch.no_grad():
            for p in net.parameters():
                p.zero_()
        plain = hold_loss()
        with torch.no_grad():
            for p, q in zip(net.parameters(), saved):
                p.copy_(q)
        log(f"unroll: held-out loss without corrections {plain:.4e}")
    log(f"unroll: K={K} J={J} crop {crop} -> {crop_out}, batch {batch}, held-out loss {history[0][2]:.4e}")
    t0 = time.perf_counter()
    run = []
    it = skipped = 0
    while it < steps:
        for g in opt.param_groups:
            g["lr"] = lr * 0.5 * (1.0 + math.cos(math.pi * it / steps))
        loss = rollout_loss(net, T, pick(rng, T.shape[0]), K, J, crop, dt, dx, gamma, th, mo, norm, whole)
        opt.zero_grad(set_to_none=True)
        loss.backward()
        norm = torch.nn.utils.clip_grad_norm_(net.parameters(), 1.0)
        if bool(torch.isfinite(norm)) and bool(torch.isfinite(loss)):
            opt.step()
            run.append(loss.detach())
        else:
            skipped += 1                      # a sample whose rollout broke down: no update
        if it + 1 == 10:
            float(loss.detach())              # waits for the device
            rate = (time.perf_counter() - t0) / 10
            log(f"unroll: {rate:.3f} s per step")
            fit = int(seconds / rate) if seconds > 0 else steps
            if fit < steps:
                log(f"unroll: {fit} steps fit in {seconds:.0f} s (of {steps} asked)")
                steps = max(fit, 10)
        if (it + 1) % log_every == 0 or it + 1 == steps:
            mean = float(torch.stack(run).mean()) if run else float("nan")
            run = []
            history.append((it + 1, mean, hold_loss()))
            log(f"unroll: step {it + 1}/{steps} loss {mean:.4e} held-out {history[-1][2]:.4e} "
                f"skipped {skipped} ({time.perf_counter() - t0:.1f} s)")
        it += 1
    return net, history


def main(argv=None):
    from . import learned as Lm
    from . import native
    ap = argparse.ArgumentParser(prog="python -m hyperflow_amd.learned_unroll")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--init", default=Lm.WEIGHTS, help="starting parameters, normalisation and fallback settings")
    t.add_argument("--out", required=True)
    t.add_argument("--seeds", type=int, default=40, help="seeds 0..n-1; the last --holdout-seeds are held out")
    t.add_argument("--holdout-seeds", type=int, default=2)
    t.add_argument("--every", type=int, default=8, help="coarse steps K between corrections")
    t.add_argument("--unroll", type=int, default=3, help="corrections J per sample")
    t.add_argument("--count", type=int, default=25, help="states per seed after the initial one")
    t.add_argument("--first", type=int, default=0, help="first start index of a sample")
    t.add_argument("--crop", type=int, default=128, help="side of the loss region of the last correction")
    t.add_argument("--batch", type=int, default=8)
    t.add_argument("--steps", type=int, default=2000)
    t.add_argument("--lr", type=float, default=1e-4)
    t.add_argument("--seconds", type=float, default=0.0, help="time budget of the training loop (0: none)")
    t.add_argument("--log-every", type=int, default=20)
    t.add_argument("--norm", choices=("l1", "l2"), default="l1", help="error norm of the loss")
    t.add_argument("--whole", action="store_true", help="whole-domain samples with the solver's boundaries")
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    Lm.ensure_tuned((Lm.COARSE, Lm.FINE), log=log)
    seeds = list(range(args.seeds))
    tr, ho = seeds[:len(seeds) - args.holdout_seeds], seeds[len(seeds) - args.holdout_seeds:]
    T = make_sequences(tr, args.every, args.count, log=log)
    H = make_sequences(ho, args.every, args.count, log=log) if ho else None
    tnet, history = train(T, args.init, args.every, args.unroll, args.crop, args.batch, args.steps, args.lr,
                          first=args.first, log=log, hold=H, seconds=args.seconds, log_every=args.log_every,
                          norm=args.norm, whole=args.whole)
    z = np.load(args.init, allow_pickle=False)
    net = native.Net(in_mean=z["in_mean"], in_scale=z["in_scale"], out_scale=z["out_scale"], gamma=float(z["gamma"]),
                     shock_threshold=float(z["shock_threshold"]), max_output=float(z["max_output"]))
    net.set_params(tnet.get_params())
    meta = dict(Lm.load_meta(args.init))
    meta["every"] = args.every          # runs correct at the interval the network was trained for
    meta["trained_every"] = args.every
    meta.setdefault("unrolled", []).append({
        "init": os.path.basename(args.init), "every": args.every, "unroll": args.unroll,
        "train_seeds": [tr[0], tr[-1]], "holdout_seeds": ho, "states_per_seed": args.count + 1, "first": args.first,
        "crop": args.crop, "batch": args.batch, "norm": args.norm, "whole": args.whole, "steps": history[-1][0],
        "lr": args.lr, "history": history})
    Lm.save(net, args.out, meta)
    log("unroll: saved " + args.out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
