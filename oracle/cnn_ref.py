"""CPU reference of the learned-correction network -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline may use this
module, and only to check the GPU kernels.  The network has no counterpart in
the reference code base; it is pinned here by independent means:

* conv() against scipy.signal.correlate2d (an unrelated implementation),
* gradients() against central finite differences of loss() in float64,
* the BF16 rounding against a direct bit-level definition (round to nearest
  even on the upper 16 bits of the IEEE single).

Conventions (the GPU kernels follow the same ones):
  parameters, layer by layer: W[9][ci][co] (tap = 3*(dy+1) + (dx+1)), b[co];
  conv:    z[co][y][x] = b[co] + sum_t sum_ci a[ci][y+dy][x+dx] * W[t][ci][co],
           zero outside the image ("same" padding);
  network: x = (u - in_mean) * in_scale; four conv+ReLU layers; a last conv;
  loss:    mean over cells and the 4 variables of (y - (target - u)/out_scale)^2.
With emulate=True the values the GPU stores in BF16 (input, weights, hidden
activations, loss gradient, hidden gradients) are rounded the same way.
"""
from __future__ import annotations

import numpy as np

LAYERS = [(4, 32), (32, 32), (32, 32), (32, 32), (32, 4)]
NPARAM = sum(9 * ci * co + co for ci, co in LAYERS)
TAPS = [(t // 3 - 1, t % 3 - 1) for t in range(9)]


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to BF16 (nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)
    return np.where(np.isnan(f), np.nan, out)


def split(params: np.ndarray):
    """[(W[9][ci][co], b[co]), ...] views of a flat parameter vector."""
    out, off = [], 0
    for ci, co in LAYERS:
        W = params[off:off + 9 * ci * co].reshape(9, ci, co)
        off += 9 * ci * co
        b = params[off:off + co]
        off += co
        out.append((W, b))
    return out


def conv(a: np.ndarray, W: np.ndarray, b=None) -> np.ndarray:
    """a[ci][ny][nx], W[9][ci][co] -> z[co][ny][nx] (float64)."""
    ci, ny, nx = a.shape
    ap = np.zeros((ci, ny + 2, nx + 2))
    ap[:, 1:-1, 1:-1] = a
    z = np.zeros((W.shape[2], ny, nx))
    for t, (dy, dx) in enumerate(TAPS):
        sh = ap[:, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]
        z += np.tensordot(W[t], sh, axes=([0], [0]))
    if b is not None:
        z += np.asarray(b, dtype=np.float64)[:, None, None]
    return z


def conv_data_grad(g: np.ndarray, W: np.ndarray) -> np.ndarray:
    """Gradient w.r.t. a of sum(g * conv(a, W)): ga[ci][q] = sum_t sum_co g[co][q - t] W[t][ci][co]."""
    co, ny, nx = g.shape
    gp = np.zeros((co, ny + 2, nx + 2))
    gp[:, 1:-1, 1:-1] = g
    ga = np.zeros((W.shape[1], ny, nx))
    for t, (dy, dx) in enumerate(TAPS):
        sh = gp[:, 1 - dy:1 - dy + ny, 1 - dx:1 - dx + nx]
        ga += np.tensordot(W[t], sh, axes=([1], [0]))
    return ga


def conv_weight_grad(a: np.ndarray, g: np.ndarray) -> np.ndarray:
    """Gradient w.r.t. W of sum(g * conv(a, W)): gW[t][ci][co] = sum_p a[ci][p + t] g[co][p]."""
    ci, ny, nx = a.shape
    ap = np.zeros((ci, ny + 2, nx + 2))
    ap[:, 1:-1, 1:-1] = a
    gW = np.zeros((9, ci, g.shape[0]))
    for t, (dy, dx) in enumerate(TAPS):
        sh = ap[:, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]
        gW[t] = np.tensordot(sh, g, axes=([1, 2], [1, 2]))
    return gW


def _r(x, emulate):
    return bf16(x) if emulate else x


def forward(params, u, in_mean, in_scale, emulate=True):
    """Network output y[4][ny][nx] and the cached activations."""
    layers = split(np.asarray(params, dtype=np.float64))
    u = np.asarray(u, dtype=np.float64)
    x = _r((u - np.asarray(in_mean, dtype=np.float64)[:, None, None]) * np.asarray(in_scale, dtype=np.float64)[:, None, None], emulate)
    acts = [x]
    a = x
    for l in range(4):
        W, b = layers[l]
        z = conv(a, _r(W, emulate), b)
        a = _r(np.maximum(z, 0.0), emulate)
        acts.append(a)
    W, b = layers[4]
    y = conv(a, _r(W, emulate), b)
    return y, acts


def loss(params, u, target, in_mean, in_scale, out_scale, emulate=False):
    y, _ = forward(params, u, in_mean, in_scale, emulate)
    t = (np.asarray(target, np.float64) - np.asarray(u, np.float64)) / np.asarray(out_scale, np.float64)[:, None, None]
    return float(np.mean((y - t) ** 2))


def gradients(params, u, target, in_mean, in_scale, out_scale, emulate=True):
    """(loss, flat gradient) of the mean squared error."""
    params = np.asarray(params, dtype=np.float64)
    layers = split(params)
    y, acts = forward(params, u, in_mean, in_scale, emulate)
    t = (np.asarray(target, np.float64) - np.asarray(u, np.float64)) / np.asarray(out_scale, np.float64)[:, None, None]
    d = y - t
    n = d.size
    L = float(np.mean(d ** 2))
    g = _r(2.0 * d / n, emulate)
    grads = [None] * 5
    for l in range(4, -1, -1):
        W, _ = layers[l]
        a_in = acts[l]
        grads[l] = (conv_weight_grad(a_in, g), g.sum(axis=(1, 2)))
        if l > 0:
            ga = conv_data_grad(g, _r(W, emulate))
            g = _r(ga * (acts[l] > 0), emulate)
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return L, flat


def pressure(U, gamma):
    """Ideal-gas pressure, evaluated in the dtype of U."""
    r, mu, mv, E = U
    half, one = U.dtype.type(0.5), U.dtype.type(1.0)
    return (U.dtype.type(gamma) - one) * (E - half * (mu * mu + mv * mv) / r)


def shock_sensor(U, gamma, theta, dtype=np.float64):
    """Per-cell flag: normalised pressure second difference above theta in x
    or y (neighbours clamped at the edges), evaluated in dtype."""
    p = pressure(np.asarray(U, dtype=dtype), gamma)
    pl = np.concatenate([p[:, :1], p[:, :-1]], axis=1)
    pr = np.concatenate([p[:, 1:], p[:, -1:]], axis=1)
    pd = np.concatenate([p[:1, :], p[:-1, :]], axis=0)
    pu = np.concatenate([p[1:, :], p[-1:, :]], axis=0)
    two = p.dtype.type(2.0)
    sx = np.abs(pr - two * p + pl) / (pr + two * p + pl)
    sy = np.abs(pu - two * p + pd) / (pu + two * p + pd)
    th = p.dtype.type(theta)
    return (sx > th) | (sy > th) | np.isnan(sx) | np.isnan(sy)


def apply_correction(U, y, out_scale, gamma, theta, dtype=np.float64, max_output=0.0):
    """Corrected state and the fallback mask: the plain state stands where the
    corrected density or pressure is not positive (or not finite), the shock
    sensor fires on the plain state, or some |y| exceeds max_output (when
    positive).  Evaluated in dtype."""
    U = np.asarray(U, dtype=dtype)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        n = U + np.asarray(out_scale, dtype)[:, None, None] * np.asarray(y, dtype)
        p = pressure(n, gamma)
    bad = ~(n[0] > 0) | ~(p > 0) | ~np.isfinite(p) | ~np.all(np.isfinite(n[1:]), axis=0)
    if theta > 0:
        bad |= shock_sensor(U, gamma, theta, dtype)
    if max_output > 0:
        bad |= np.any(np.abs(np.asarray(y, dtype)) > dtype(max_output), axis=0)
    out = np.where(bad[None], U, n)
    return out, bad


def restrict2(U):
    """2x2 block average (conservative restriction) of planes [4][ny][nx]."""
    U = np.asarray(U, dtype=np.float64)
    return 0.25 * ((U[:, 0::2, 0::2] + U[:, 0::2, 1::2]) + (U[:, 1::2, 0::2] + U[:, 1::2, 1::2]))
