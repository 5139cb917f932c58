"""FP64 checker for the surrogate-gradient adjoint (test infrastructure only).

An FP64 torch-autograd unroll of the same recurrent LIF network on the CPU, used to
check the HIP adjoint's gradients, plus central finite differences of its loss.
The differentiable form has no counterpart in the reference, so this unroll is the
reference: it is pinned in tests/test_adjoint_oracle.py against finite differences
(surrogate off: the exact gradient of the piecewise-smooth loss) and against closed
forms of the LIF step (RK4 factor) and of the leaky readout.

Model (per step t, state [N, B]):
    I  = W_in x_t + W_rec s_{t-1}
    Vm = g V + (1 - g)(V_rest + R I),  g = 1 - h + h^2/2 - h^3/6 + h^4/24, h = dt/tau_m
    s  = H(Vm - V_th)   (backward: fast sigmoid 1/(1 + slope|Vm - V_th|)^2 in mode 1, 0 in mode 0);
         mode 2: s = 1/2 + u/(2(1 + |u|)), u = slope (Vm - V_th), differentiated exactly
    V  = Vm - (V_th - V_reset) s
    y  = beta y + (1 - beta) W_out s,  beta = exp(-dt/tau_out)
    loss = mean((y - target)^2)
"""
from __future__ import annotations

import math

import torch


class _Spike(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, forced, slope, mode):
        ctx.save_for_backward(x)
        ctx.slope, ctx.mode = slope, mode
        return forced.clone() if forced is not None else (x > 0).to(x.dtype)

    @staticmethod
    def backward(ctx, g):
        (x,) = ctx.saved_tensors
        if ctx.mode == 0:
            return torch.zeros_like(g), None, None, None
        return g / (1.0 + ctx.slope * x.abs()) ** 2, None, None, None


def spike(x, p, forced=None):
    if p.mode == 2:
        u = p.slope * x
        return 0.5 + 0.5 * u / (1.0 + u.abs())
    return _Spike.apply(x, forced, p.slope, p.mode)


def rk4_factor(dt: float, tau: float) -> float:
    h = dt / tau
    return 1.0 - h + h * h / 2.0 - h ** 3 / 6.0 + h ** 4 / 24.0


def run(W_in, W_rec, W_out, x, p, target=None, spikes=None):
    """FP64 forward. x [T, C, B]; spikes (optional [T, N, B]) replays a given raster in the
    forward while the surrogate is still evaluated at the FP64 membrane potential.
    Returns (y [T, M, B], raster [T, N, B], loss or None)."""
    T, _, B = x.shape
    N = W_rec.shape[0]
    g = rk4_factor(p.dt, p.tau_m)
    beta = math.exp(-p.dt / p.tau_out)
    V = torch.full((N, B), p.v_rest, dtype=torch.float64)
    s_prev = torch.zeros((N, B), dtype=torch.float64)
    y = torch.zeros((W_out.shape[0], B), dtype=torch.float64)
    ys, ss = [], []
    for t in range(T):
        I = W_in @ x[t] + W_rec @ s_prev
        Vm = g * V + (1.0 - g) * (p.v_rest + p.r * I)
        s = spike(Vm - p.v_th, p, None if spikes is None else spikes[t])
        V = Vm - (p.v_th - p.v_reset) * s
        y = beta * y + (1.0 - beta) * (W_out @ s)
        ys.append(y)
        ss.append(s.detach())
        s_prev = s
    Y = torch.stack(ys)
    loss = None
    if target is not None:
        loss = ((Y - target) ** 2).mean()
    return Y, torch.stack(ss), loss


def gradients(W_in, W_rec, W_out, x, p, target, spikes=None):
    """FP64 autograd gradients (d loss / d W_in, W_rec, W_out) and the loss."""
    ws = [w.detach().to(torch.float64).clone().requires_grad_(True) for w in (W_in, W_rec, W_out)]
    _, raster, loss = run(ws[0], ws[1], ws[2], x.to(torch.float64), p, target.to(torch.float64),
                          None if spikes is None else spikes.to(torch.float64))
    loss.backward()
    return [w.grad for w in ws], float(loss.detach()), raster


def finite_difference(W_in, W_rec, W_out, x, p, target, which: int, idx, eps: float = 1e-6):
    """Central difference of the FP64 loss w.r.t. weight tensor `which` at flat indices `idx`.
    Returns (fd values, True if no perturbation changed the spike raster; always True in mode 2)."""
    base = [w.detach().to(torch.float64).clone() for w in (W_in, W_rec, W_out)]
    x64, tg = x.to(torch.float64), target.to(torch.float64)
    _, r0, _ = run(*base, x64, p, tg)
    out, stable = [], True
    with torch.no_grad():
        for i in idx:
            vals = []
            for sgn in (1.0, -1.0):
                w = [b.clone() for b in base]
                w[which].view(-1)[i] += sgn * eps
                _, r, l = run(*w, x64, p, tg)
                stable &= p.mode == 2 or bool(torch.equal(r, r0))
                vals.append(float(l))
            out.append((vals[0] - vals[1]) / (2 * eps))
    return torch.tensor(out, dtype=torch.float64), stable
