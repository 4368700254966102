"""Pins of the CPU reference of the learned-correction network (oracle/cnn_ref.py).

The network has no counterpart in the reference code base, so the reference
implementation is checked against independent routes: scipy's correlation,
adjoint identities, central finite differences, and bit-level BF16 rounding.
"""
import numpy as np
import pytest
from scipy.signal import correlate2d

import cnn_ref as R


def test_bf16_rounding_bits():
    def f(bits):
        return np.array([bits], np.uint32).view(np.float32)

    # exact, ties to even (down and up), above the tie, negative, NaN
    assert R.bf16(f(0x3F800000))[0] == 1.0
    assert R.bf16(f(0x3F808000)).astype(np.float32).view(np.uint32)[0] == 0x3F800000
    assert R.bf16(f(0x3F818000)).astype(np.float32).view(np.uint32)[0] == 0x3F820000
    assert R.bf16(f(0x3F808001)).astype(np.float32).view(np.uint32)[0] == 0x3F810000
    assert R.bf16(f(0xBF808001)).astype(np.float32).view(np.uint32)[0] == 0xBF810000
    assert np.isnan(R.bf16(np.array([np.nan], np.float32))[0])
    x = np.random.default_rng(0).normal(size=1000).astype(np.float32)
    r = R.bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8)
    assert np.all(r.astype(np.float32).view(np.uint32) & 0xFFFF == 0)


def test_conv_matches_scipy_correlation():
    rng = np.random.default_rng(1)
    a = rng.normal(size=(3, 7, 9))
    W = rng.normal(size=(9, 3, 5))
    b = rng.normal(size=5)
    z = R.conv(a, W, b)
    for co in range(5):
        ref = np.full((7, 9), b[co])
        for ci in range(3):
            K = np.zeros((3, 3))
            for t, (dy, dx) in enumerate(R.TAPS):
                K[dy + 1, dx + 1] = W[t, ci, co]
            ref += correlate2d(a[ci], K, mode="same", boundary="fill", fillvalue=0.0)
        np.testing.assert_allclose(z[co], ref, rtol=1e-12, atol=1e-12)


def test_conv_adjoint_identities():
    rng = np.random.default_rng(2)
    a = rng.normal(size=(4, 6, 10))
    W = rng.normal(size=(9, 4, 3))
    g = rng.normal(size=(3, 6, 10))
    lhs = np.sum(g * R.conv(a, W))
    assert np.isclose(lhs, np.sum(a * R.conv_data_grad(g, W)), rtol=1e-12)
    assert np.isclose(lhs, np.sum(W * R.conv_weight_grad(a, g)), rtol=1e-12)


def _problem(seed, ny=6, nx=8):
    rng = np.random.default_rng(seed)
    params = np.concatenate([np.concatenate([rng.normal(scale=np.sqrt(2.0 / (9 * ci)), size=9 * ci * co),
                                             rng.normal(scale=0.1, size=co)]) for ci, co in R.LAYERS])
    u = np.stack([rng.uniform(0.5, 2, (ny, nx)), rng.normal(size=(ny, nx)), rng.normal(size=(ny, nx)),
                  rng.uniform(2, 4, (ny, nx))])
    target = u + 0.05 * rng.normal(size=u.shape)
    norm = dict(in_mean=[1.0, 0.0, 0.0, 3.0], in_scale=[2.0, 1.0, 1.0, 1.0], out_scale=[0.05, 0.05, 0.05, 0.05])
    return params, u, target, norm


def test_gradients_match_finite_differences():
    params, u, target, norm = _problem(3)
    L, g = R.gradients(params, u, target, emulate=False, **norm)
    assert np.isclose(L, R.loss(params, u, target, **norm), rtol=1e-14)
    rng = np.random.default_rng(4)
    offs = np.cumsum([0] + [9 * ci * co + co for ci, co in R.LAYERS])
    idx = list(rng.choice(R.NPARAM, 40, replace=False))
    idx += [int(offs[l + 1] - 1) for l in range(5)]   # a bias of every layer
    h = 1e-6
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        fd = (R.loss(p1, u, target, **norm) - R.loss(p2, u, target, **norm)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(g[i])) + 1e-9, (i, fd, g[i])


def test_emulated_gradients_close_to_exact():
    params, u, target, norm = _problem(5, 8, 16)
    _, g64 = R.gradients(params, u, target, emulate=False, **norm)
    _, g16 = R.gradients(params, u, target, emulate=True, **norm)
    assert np.linalg.norm(g16 - g64) <= 0.05 * np.linalg.norm(g64)


def test_restriction_is_conservative():
    rng = np.random.default_rng(6)
    U = rng.uniform(0.1, 2.0, size=(4, 10, 14))
    c = R.restrict2(U)
    assert c.shape == (4, 5, 7)
    np.testing.assert_allclose(4.0 * c.sum(axis=(1, 2)), U.sum(axis=(1, 2)), rtol=1e-14)
    np.testing.assert_allclose(c[:, 2, 3], U[:, 4:6, 6:8].mean(axis=(1, 2)), rtol=1e-15)


def test_fallback_on_injected_bad_outputs():
    gamma = 1.4
    ny, nx = 5, 6
    W = np.array([1.0, 0.2, -0.1, 1.0])                       # rho, u, v, p
    U = np.empty((4, ny, nx))
    U[0], U[1], U[2] = W[0], W[0] * W[1], W[0] * W[2]
    U[3] = W[3] / (gamma - 1) + 0.5 * W[0] * (W[1] ** 2 + W[2] ** 2)
    y = np.zeros_like(U)
    y[0, 1, 2] = -2.0          # density -> -1
    y[3, 3, 4] = -10.0         # pressure -> negative
    y[1, 0, 0] = np.inf        # non-finite
    y[0, 4, 5] = 0.5           # fine
    out, bad = R.apply_correction(U, y, [1.0] * 4, gamma, 0.0)
    assert bad[1, 2] and bad[3, 4] and bad[0, 0] and not bad[4, 5]
    assert bad.sum() == 3
    np.testing.assert_array_equal(out[:, bad], U[:, bad])
    assert out[0, 4, 5] == U[0, 4, 5] + 0.5


def test_shock_sensor_flags_pressure_jump_only():
    gamma = 1.4
    ny, nx = 4, 10
    rho = np.where(np.arange(nx) < 3, 1.0, 0.125)[None, :].repeat(ny, 0)     # contact: pressure continuous
    p = np.where(np.arange(nx) < 7, 1.0, 10.0)[None, :].repeat(ny, 0)        # shock between x = 6 and 7
    U = np.stack([rho, 0 * rho, 0 * rho, p / (gamma - 1)])
    # second differences: 9/13 at x = 6, 9/31 at x = 7, 0 elsewhere
    s = R.shock_sensor(U, gamma, 0.25)
    assert s[:, 6].all() and s[:, 7].all()
    assert not s[:, :6].any() and not s[:, 8:].any()
    s = R.shock_sensor(U, gamma, 0.3)
    assert s[:, 6].all() and not s[:, 7].any()
