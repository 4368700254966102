"""The differentiable step and network of the solver-in-the-loop trainer
(hyperflow_amd/learned_unroll.py) against the oracle: the step in float64
against the oracle's MUSCL-Hancock residual (itself bitwise against the
reference), the network and the correction rules against oracle/cnn_ref.py
(pinned by scipy, finite differences and bit-level BF16 rounding)."""
import numpy as np
import pytest
import torch

import cnn_ref
import hfo
from hyperflow_amd import case as C
from hyperflow_amd import learned_unroll as LU


def _oracle_step(case, fields, warm_calls=0):
    """(padded state, step dt, predictor dt, oracle result on the interior)."""
    m = hfo.OracleMesh(case, fields)
    for _ in range(warm_calls):
        m.time_step()
    m.apply_bcs()
    full = m.fields(0)
    dt = m.time_step()
    L = m.residual(0)
    dt_pred = m.dt_log()[-1]
    return full, dt, dt_pred, (dt * L + full)[:, 2:-2, 2:-2]


@pytest.mark.parametrize("limiter", ["VanLeer", "MinBee"])
@pytest.mark.parametrize("seed", [5, 6])
def test_step_matches_oracle(limiter, seed):
    case = C.quad_case(26, 31, seed, scheme="MUSCLHancock", limiter=limiter, cfl=0.5)
    full, dt, dt_pred, want = _oracle_step(case, C.perturbed_fields(case, seed))
    b = case.blocks[0]
    got = LU.muscl_step(torch.from_numpy(full), dt, b.dx, case.gamma, dt_pred, limiter.lower(), dy=b.dy).numpy()
    assert got.shape == want.shape
    assert np.abs(got - want).max() <= 1e-14 * np.abs(want).max()


def test_step_uses_the_predictor_dt_in_the_predictor():
    """Nine earlier TimeStep calls: the step's dt is the 10th call (0.2 CFL),
    the predictor's the 11th (full CFL), five times larger."""
    case = C.quad_case(24, 24, 7, scheme="MUSCLHancock", limiter="VanLeer", cfl=0.5)
    full, dt, dt_pred, want = _oracle_step(case, C.perturbed_fields(case, 7), warm_calls=9)
    assert dt_pred == pytest.approx(5 * dt, rel=0.2)
    b = case.blocks[0]
    U = torch.from_numpy(full)
    good = LU.muscl_step(U, dt, b.dx, case.gamma, dt_pred, dy=b.dy).numpy()
    swapped = LU.muscl_step(U, dt, b.dx, case.gamma, dt, dy=b.dy).numpy()
    scale = np.abs(want).max()
    assert np.abs(good - want).max() <= 1e-14 * scale
    assert np.abs(swapped - want).max() > 1e-8 * scale


def _net(rng, scale=0.15):
    params = rng.normal(0.0, scale, cnn_ref.NPARAM)
    mean, iscale, oscale = rng.normal(0, 0.5, 4), rng.uniform(0.5, 2, 4), rng.uniform(1e-3, 1e-2, 4)
    net = LU.TorchNet(mean, iscale, oscale).double()
    net.set_params(params)
    return net, params, mean, iscale, oscale


def test_param_layout_round_trip():
    net, params, *_ = _net(np.random.default_rng(0))
    assert np.array_equal(net.get_params(), params.astype(np.float32))


def test_network_matches_cnn_ref():
    rng = np.random.default_rng(1)
    net, params, mean, iscale, _ = _net(rng)
    u = rng.normal(1.0, 0.3, (4, 22, 27))
    want = cnn_ref.forward(params, u, mean, iscale, emulate=False)[0][:, 5:-5, 5:-5]
    got = net(torch.from_numpy(u)).detach().numpy()
    assert got.shape == want.shape
    assert np.abs(got - want).max() <= 1e-12 * max(1.0, np.abs(want).max())


def test_correction_rules_match_cnn_ref():
    """Fallbacks: a pressure jump (sensor), outputs beyond the bound, and
    corrections that make the density negative."""
    rng = np.random.default_rng(2)
    net, params, mean, iscale, oscale = _net(rng, scale=0.1)
    oscale = np.array([0.05, 0.05, 0.05, 0.05])
    net.out_scale.copy_(torch.from_numpy(oscale).view(4, 1, 1))
    ny, nx = 30, 34
    rho = rng.uniform(0.1, 1.0, (ny, nx))
    u, v = rng.normal(0, 0.3, (2, ny, nx))
    p = np.where(np.arange(nx)[None, :] < 15, 1.0, 0.1) * rng.uniform(0.9, 1.1, (ny, nx))
    U = np.stack([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)])
    y = cnn_ref.forward(params, U, mean, iscale, emulate=False)[0]
    for th, mo in ((0.5, 1.5), (0.5, 0.0), (0.0, 2.0), (0.0, 0.0)):
        want, bad = cnn_ref.apply_correction(U, y, oscale, 1.4, th, max_output=mo)
        got, gbad = LU.correct(net, torch.from_numpy(U), 1.4, th, mo)
        inner = (slice(None), slice(5, -5), slice(5, -5))
        assert np.array_equal(gbad.numpy(), bad[5:-5, 5:-5])
        assert 0 < gbad.sum() < gbad.numel()
        assert np.abs(got.detach().numpy() - want[inner]).max() <= 1e-12


def test_rollout_regions_line_up():
    """With a zero network the rollout of a crop equals the plain steps of the
    whole field on the same cells, so a sequence made by plain steps gives a
    loss of exactly zero; a sequence shifted by one cell does not."""
    rng = np.random.default_rng(3)
    K, J, n, crop = 2, 2, 64, 40
    case = C.quad_case(n, n, 3, scheme="MUSCLHancock", limiter="VanLeer", cfl=0.5)
    g = case.ghost
    U0 = torch.from_numpy(C.perturbed_fields(case, 3)[0][:, g:-g, g:-g].copy())
    dt, dx = 0.1 / n, 1.0 / n
    T = [U0]
    for _ in range(J + 1):
        U = T[-1]
        for _ in range(K):
            U = LU.muscl_step(torch.nn.functional.pad(U[None], (2, 2, 2, 2), mode="replicate")[0], dt, dx)
        T.append(U)
    T = torch.stack(T)[None]
    net, *_ = _net(rng)
    net.set_params(np.zeros(cnn_ref.NPARAM))
    picks = [(0, 0, 5, 11), (0, 1, 20, 3)]
    assert float(LU.rollout_loss(net, T, picks, K, J, crop, dt, dx, 1.4, 0.5, 8.0)) == 0.0
    shifted = [(0, 0, 5, 12)]
    T2 = torch.cat([T[:, :1], T[:, 1:].roll(1, dims=-1)], dim=1)
    assert float(LU.rollout_loss(net, T2, shifted, K, J, crop, dt, dx, 1.4, 0.5, 8.0)) > 0.0


def test_gradient_through_steps_and_correction_matches_finite_differences():
    """Directional derivative of the two-step, one-correction rollout loss
    with respect to the parameters, autograd against central differences
    (float64, smooth state, so no limiter or Riemann branch changes within
    the difference)."""
    rng = np.random.default_rng(4)
    net, params, *_ = _net(rng, scale=0.05)
    net.out_scale.copy_(torch.full((4, 1, 1), 1e-3, dtype=torch.float64))
    n = 40
    x = np.linspace(0, 1, n)
    X, Y = np.meshgrid(x, x)
    rho = 1 + 0.2 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y)
    U = np.stack([rho, 0.3 * rho, -0.2 * rho, 2.5 + 0.1 * np.cos(2 * np.pi * (X + Y))])
    T = torch.from_numpy(np.stack([U, U + 1e-3 * rng.normal(size=U.shape)]))[None]
    picks = [(0, 0, 0, 0)]
    loss = LU.rollout_loss(net, T, picks, 2, 1, n, 0.2 / n, 1.0 / n, 1.4, 0.0, 0.0)
    net.zero_grad()
    loss.backward()
    grad = np.concatenate([np.concatenate([c.weight.grad.numpy().transpose(2, 3, 1, 0).reshape(-1),
                                           c.bias.grad.numpy()]) for c in net.convs])
    d = np.random.default_rng(9).normal(size=params.size)
    eps = 1e-7          # small enough that no ReLU changes sign within the difference

    def at(p):
        net.set_params(p)
        with torch.no_grad():
            return float(LU.rollout_loss(net, T, picks, 2, 1, n, 0.2 / n, 1.0 / n, 1.4, 0.0, 0.0))

    fd = (at(params + eps * d) - at(params - eps * d)) / (2 * eps)
    assert np.dot(grad, d) == pytest.approx(fd, rel=1e-5)


def test_training_lowers_the_rollout_loss(tmp_path):
    """A few Adam steps on a synthetic sequence whose states run ahead of the
    plain coarse steps by a fixed increment: the held-out rollout loss drops."""
    torch.manual_seed(0)
    rng = np.random.default_rng(5)
    K, J, n = 1, 2, 48
    case = C.quad_case(n, n, 5, scheme="MUSCLHancock", limiter="VanLeer", cfl=0.5)
    g = case.ghost
    from hyperflow_amd import learned as Lm
    dt, dx = Lm.dt_for(Lm.COARSE), 1.0 / Lm.COARSE
    shift = torch.tensor([1e-3, 0.0, 0.0, 2e-3], dtype=torch.float64).view(4, 1, 1)
    seqs = []
    for s in (5, 6):
        U = torch.from_numpy(C.perturbed_fields(case, s)[0][:, g:-g, g:-g].copy())
        T = [U]
        for _ in range(5):
            V = T[-1]
            for _ in range(K):
                V = LU.muscl_step(torch.nn.functional.pad(V[None], (2, 2, 2, 2), mode="replicate")[0], dt, dx)
            T.append(V + shift)
        seqs.append(torch.stack(T))
    T = torch.stack(seqs).float()
    init = tmp_path / "init.npz"
    np.savez(init, params=rng.normal(0, 0.02, cnn_ref.NPARAM).astype(np.float32),
             in_mean=np.array([0.5, 0, 0, 1.5], np.float32), in_scale=np.ones(4, np.float32),
             out_scale=np.full(4, 1e-3, np.float32), gamma=np.float32(1.4), shock_threshold=np.float32(0.5),
             max_output=np.float32(8.0), meta=np.array("{}"))
    net, history = LU.train(T[:1], str(init), K, J, crop_out=8, batch=2, steps=60, lr=3e-3, log=lambda m: None, norm="l2",
                            hold=T[1:], log_every=20)
    assert history[-1][2] < 0.5 * history[0][2]


def test_nan_wave_speed_gives_zero_flux_and_finite_gradient():
    """A face state with negative pressure: the flux is zero, as in the
    kernel (hf_physics.h f32::hllc), and the gradient stays finite."""
    L = torch.tensor([1.0, 0.2, 0.1, -0.5], dtype=torch.float64).view(1, 4, 1, 1).repeat(1, 1, 1, 2)
    R = torch.tensor([0.8, 0.1, 0.0, 2.0], dtype=torch.float64).view(1, 4, 1, 1).repeat(1, 1, 1, 2)
    R[0, :, 0, 1] = torch.tensor([0.9, 0.0, 0.1, 2.2])
    L[0, 3, 0, 1] = 2.5               # the second face is valid
    L.requires_grad_()
    F = LU._hllc(True, L, R, 1.4, 0.4)
    assert torch.all(F[..., 0] == 0) and torch.all(F[..., 1] != 0)
    F.sum().backward()
    assert torch.isfinite(L.grad).all()


def test_transmissive_step_matches_oracle():
    """The whole-domain step (replicate padding) against the oracle's step
    with transmissive boundaries."""
    case = C.quad_case(30, 30, 8, scheme="MUSCLHancock", limiter="VanLeer", cfl=0.5)
    full, dt, dt_pred, want = _oracle_step(case, C.perturbed_fields(case, 8))
    assert dt == dt_pred
    b = case.blocks[0]
    inner = torch.from_numpy(full[:, 2:-2, 2:-2].copy())[None]
    got = LU.transmissive_step(inner, dt, b.dx, case.gamma)[0].numpy()
    assert np.abs(got - want).max() <= 1e-14 * np.abs(want).max()


def test_whole_domain_correction_matches_cnn_ref():
    """Zero padding in the network and clamped sensor neighbours at the
    edges, on every cell of the field."""
    rng = np.random.default_rng(12)
    net, params, mean, iscale, _ = _net(rng, scale=0.1)
    oscale = np.full(4, 0.05)
    net.out_scale.copy_(torch.from_numpy(oscale).view(4, 1, 1))
    ny, nx = 21, 26
    rho = rng.uniform(0.1, 1.0, (ny, nx))
    u, v = rng.normal(0, 0.3, (2, ny, nx))
    p = np.where(np.arange(nx)[None, :] < 11, 1.0, 0.1) * rng.uniform(0.9, 1.1, (ny, nx))
    U = np.stack([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)])
    y = cnn_ref.forward(params, U, mean, iscale, emulate=False)[0]
    want, bad = cnn_ref.apply_correction(U, y, oscale, 1.4, 0.5, max_output=1.5)
    got, gbad = LU.correct(net, torch.from_numpy(U)[None], 1.4, 0.5, 1.5, whole=True)
    assert np.array_equal(gbad[0].numpy(), bad) and 0 < bad.sum() < bad.size
    assert np.abs(got[0].detach().numpy() - want).max() <= 1e-12
