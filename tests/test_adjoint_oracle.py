"""Pins the FP64 adjoint checker (oracle/adjoint_ref.py): autograd of the unroll against
central finite differences, and the step/readout algebra against closed forms."""
import math

import numpy as np
import pytest
import torch

from oracle import adjoint_ref as R
from neuronbridge_amd.adjoint import LifParams


def small(n=24, b=3, T=40, n_in=10, seed=3, gain=3.0):
    rng = np.random.default_rng(seed)
    W_in = torch.tensor(rng.standard_normal((n, n_in)) * gain / np.sqrt(n_in))
    W_rec = torch.tensor(rng.standard_normal((n, n)) / np.sqrt(n))
    W_out = torch.tensor(rng.standard_normal((4, n)) / np.sqrt(n))
    x = torch.tensor((rng.random((T, n_in, b)) < 0.2).astype(np.float64))
    tgt = torch.tensor(np.sin(np.arange(T)[:, None, None] * 0.2 + np.arange(4)[None, :, None]) * np.ones((1, 1, b)))
    return W_in, W_rec, W_out, x, tgt


def test_rk4_factor_is_rk4_on_linear_ode():
    # explicit RK4 stages on dV/dt = -V/tau from V = 1 give the factor exactly, and it
    # approximates exp(-h) to fifth order
    for dt, tau in ((1.0, 20.0), (0.1, 10.0), (0.5, 3.0)):
        h = dt / tau
        f = lambda v: -v / tau
        k1 = f(1.0); k2 = f(1.0 + dt / 2 * k1); k3 = f(1.0 + dt / 2 * k2); k4 = f(1.0 + dt * k3)
        assert R.rk4_factor(dt, tau) == pytest.approx(1.0 + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4), rel=1e-15)
        assert abs(R.rk4_factor(dt, tau) - math.exp(-h)) <= h ** 5 / 100


def test_readout_closed_form():
    # one spiking neuron, unit readout weight: y_t = (1-beta) * sum_k beta^(t-k) s_k
    p = LifParams(mode=1)
    T = 30
    W_in = torch.tensor([[100.0]], dtype=torch.float64)  # drives the neuron over threshold every step
    W_rec = torch.zeros((1, 1), dtype=torch.float64)
    W_out = torch.ones((1, 1), dtype=torch.float64)
    x = torch.ones((T, 1, 1), dtype=torch.float64)
    y, s, _ = R.run(W_in, W_rec, W_out, x, p)
    beta = math.exp(-p.dt / p.tau_out)
    sv = s[:, 0, 0].numpy()
    expect = [(1 - beta) * sum(beta ** (t - k) * sv[k] for k in range(t + 1)) for t in range(T)]
    assert np.allclose(y[:, 0, 0].numpy(), expect, rtol=1e-12)
    assert sv.sum() > 5


@pytest.mark.parametrize("which", [0, 1, 2])
def test_autograd_matches_finite_differences_smooth_mode(which):
    p = LifParams(mode=2)
    W_in, W_rec, W_out, x, tgt = small()
    grads, loss, _ = R.gradients(W_in, W_rec, W_out, x, p, tgt)
    g = grads[which].reshape(-1)
    idx = np.random.default_rng(which).choice(g.numel(), size=8, replace=False)
    fd, _ = R.finite_difference(W_in, W_rec, W_out, x, p, tgt, which, idx, eps=1e-6)
    ref = g[idx]
    assert torch.max(torch.abs(fd - ref)) <= 1e-7 + 1e-5 * torch.max(torch.abs(ref))
    assert torch.max(torch.abs(ref)) > 1e-6


def test_surrogate_off_is_exact_gradient():
    # mode 0: the loss depends on W_in/W_rec only through spikes, so their gradient is 0, and
    # W_out's gradient equals finite differences (no spike flips)
    p = LifParams(mode=0)
    W_in, W_rec, W_out, x, tgt = small(seed=11)
    grads, _, raster = R.gradients(W_in, W_rec, W_out, x, p, tgt)
    assert raster.sum() > 5
    assert float(grads[0].abs().max()) == 0.0 and float(grads[1].abs().max()) == 0.0
    idx = np.arange(0, grads[2].numel(), 7)
    fd, stable = R.finite_difference(W_in, W_rec, W_out, x, p, tgt, 2, idx, eps=1e-7)
    assert stable
    assert torch.allclose(fd, grads[2].reshape(-1)[idx], rtol=1e-6, atol=1e-9)


def test_plain_autograd_unroll_matches_fp64_oracle():
    # the plain FP32 unroll timed beside the adjoint is the same network: on the oracle's spike
    # train its gradients agree with the FP64 autograd gradients to FP32 accuracy
    from neuronbridge_amd.adjoint import LifParams, autograd_loss_and_grads
    rng = np.random.default_rng(4)
    N, cin, M, B, T = 48, 12, 3, 5, 80
    Wi = rng.normal(0, 1 / np.sqrt(cin), (N, cin)).astype(np.float32)
    Wr = rng.normal(0, 1 / np.sqrt(N), (N, N)).astype(np.float32)
    Wo = rng.normal(0, 1 / np.sqrt(N), (M, N)).astype(np.float32)
    x = (rng.random((T, cin, B)) < 0.3).astype(np.float32)
    tg = rng.normal(0, 1, (T, M, B)).astype(np.float32)
    p = LifParams()
    args = [torch.tensor(a) for a in (Wi, Wr, Wo, x, tg)]
    g64, loss64, raster = R.gradients(args[0], args[1], args[2], args[3], p, args[4])
    assert 0.02 < float(raster.mean()) < 0.5
    loss32, g32, s32 = autograd_loss_and_grads(*args, p, spikes=raster.to(torch.float32))
    assert torch.equal(s32, raster.to(torch.float32))
    assert abs(float(loss32) - loss64) <= 1e-5 * abs(loss64)
    for a, b in zip(g32, g64):
        err = float((a.double() - b).norm() / b.norm())
        assert err < 1e-4, err
