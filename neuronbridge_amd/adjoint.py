"""Surrogate-gradient adjoint of a recurrent LIF network, behind torch.autograd.Function.

Forward, per step t (state [N, B] float32, B trials):
    I_t  = W_in x_t + W_rec s_{t-1}             (the input term for all t in one GEMM; the recurrent
                                                  term inside the step kernel: the rows of W_rec^T of
                                                  the neurons that fired, or MFMA for smooth spikes)
    V_t, s_t = LIF step                          (hand-written HIP kernel: RK4-exact affine update,
                                                  strict threshold, subtractive reset)
    y_t  = beta y_{t-1} + (1 - beta) W_out s_t   (GEMM for all t + HIP scan kernel)
Backward walks the steps in reverse with the hand-written adjoint step kernel
(fast-sigmoid surrogate, slope 10 by default) fused with the product W_rec^T dI_{t+1}.
Only the membrane potential at every `checkpoint`-th step is kept from the forward pass;
each segment is recomputed before its backward sweep (checkpoint=0 keeps every step).
Weight gradients are accumulated per segment as large GEMMs over (steps x trials).

The per-step recurrent products run inside libnbx's step kernels, one launch per step from a C
loop: forward with binary spikes as a sum of the W_rec^T rows of the neurons that fired
(nbx_lif_fwd_segment_sparse), with smooth spikes and backward on FP32 MFMA tiles
(nbx_lif_fwd_segment / nbx_lif_bwd_segment); the large
GEMMs over all steps (input currents, readout, weight gradients) run through torch (hipBLASLt),
all on torch's current stream.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import torch

from . import _lib as L


@dataclass
class LifParams:
    tau_m: float = 20.0      # ms
    dt: float = 1.0          # ms
    v_rest: float = 0.0
    v_th: float = 1.0
    v_reset: float = 0.0
    r: float = 20.0          # input resistance: (1 - g) * r ~ 1, so one unit of current moves V by ~1 per step
    tau_out: float = 20.0    # readout time constant, ms
    slope: float = 10.0      # fast-sigmoid surrogate slope
    mode: int = 1            # 1 surrogate gradient, 0 exact (sigma' = 0), 2 smooth spike (checking)

    @property
    def g(self) -> float:
        h = self.dt / self.tau_m
        return 1.0 - h + h * h / 2.0 - h ** 3 / 6.0 + h ** 4 / 24.0

    @property
    def beta(self) -> float:
        return math.exp(-self.dt / self.tau_out)

    def c_params(self, n: int, batch: int) -> L.LifAdjParams:
        g = self.g
        return L.LifAdjParams(n=n, batch=batch, g=g, one_minus_g=1.0 - g, r=self.r, v_rest=self.v_rest,
                              v_th=self.v_th, delta_reset=self.v_th - self.v_reset, slope=self.slope,
                              mode=self.mode)


def _ptr(t: torch.Tensor | None):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def _stream() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _fwd_segment(cp, IN, W_rec, W_rec_t, bits, s_init, V, t0, t1, S_out, Vm_out=None):
    """Forward steps [t0, t1) from state V (updated in place) and the previous step's spikes
    s_init; spikes of step t go to S_out[t - t0], membrane potentials to Vm_out[t - t0].
    One kernel per step. Binary spikes (modes 0, 1): W_rec s_{t-1} as the sum of the W_rec^T rows
    of the neurons that fired (bits: spike-word workspace); smooth spikes (mode 2): the dense
    product on the matrix cores."""
    assert S_out.is_contiguous() and (Vm_out is None or Vm_out.is_contiguous()) and s_init.is_contiguous()
    if cp.mode == 2:
        L.call("nbx_lif_fwd_segment", C.byref(cp), _ptr(W_rec), _ptr(IN[t0]), _ptr(s_init), _ptr(V), _ptr(S_out),
               _ptr(Vm_out), t1 - t0, _stream())
    else:
        L.call("nbx_lif_fwd_segment_sparse", C.byref(cp), _ptr(W_rec_t), _ptr(IN[t0]), _ptr(s_init), _ptr(V),
               _ptr(S_out), _ptr(Vm_out), _ptr(bits), t1 - t0, _stream())


def _spike_words(N: int, B: int, device) -> torch.Tensor:
    return torch.empty(2 * B * ((N + 31) // 32), device=device, dtype=torch.int32)


class LifAdjoint(torch.autograd.Function):
    """y, spikes = LifAdjoint.apply(W_in, W_rec, W_out, x, params, checkpoint)

    x: [T, C_in, B] input spikes (float32), W_in: [N, C_in], W_rec: [N, N], W_out: [M, N].
    Returns readout traces y: [T, M, B] and the spike raster [T, N, B] (not differentiable)."""

    @staticmethod
    def forward(ctx, W_in, W_rec, W_out, x, p: LifParams, checkpoint: int = 50):
        T, cin, B = x.shape
        N = W_rec.shape[0]
        M = W_out.shape[0]
        assert W_in.shape == (N, cin) and W_rec.shape == (N, N) and W_out.shape[1] == N
        for tns in (W_in, W_rec, W_out, x):
            assert tns.is_cuda and tns.dtype == torch.float32 and tns.is_contiguous()
        cp = p.c_params(N, B)
        # input currents of every step in one GEMM: IN[t] = W_in @ x[t]
        IN = torch.matmul(W_in, x.permute(1, 0, 2).reshape(cin, T * B)).reshape(N, T, B).permute(1, 0, 2).contiguous()
        S = torch.empty((T, N, B), device=x.device, dtype=torch.float32)
        V = torch.full((N, B), p.v_rest, device=x.device, dtype=torch.float32)
        K = checkpoint if checkpoint and checkpoint > 0 else T
        ckpts = []
        zero = torch.zeros_like(V)
        W_rec_t = W_rec.t().contiguous()
        bits = _spike_words(N, B, x.device)
        for t0 in range(0, T, K):
            ckpts.append(V.clone())
            _fwd_segment(cp, IN, W_rec, W_rec_t, bits, S[t0 - 1] if t0 > 0 else zero, V, t0, min(T, t0 + K), S[t0:])
        Z = torch.matmul(W_out, S)  # [T, M, B]
        y = torch.empty_like(Z)
        L.call("nbx_readout_fwd", T, M * B, p.beta, _ptr(Z), _ptr(y), C.c_void_p(0), _stream())
        ctx.save_for_backward(W_in, W_rec, W_rec_t, W_out, x, S, IN, torch.stack(ckpts))
        ctx.p = p
        ctx.K = K
        ctx.mark_non_differentiable(S)
        return y, S

    @staticmethod
    def backward(ctx, gy, _gS):
        W_in, W_rec, W_rec_t, W_out, x, S, IN, ckpts = ctx.saved_tensors
        p, K = ctx.p, ctx.K
        T, cin, B = x.shape
        N, M = W_rec.shape[0], W_out.shape[0]
        cp = p.c_params(N, B)
        gy = gy.contiguous()
        dZ = torch.empty_like(gy)
        L.call("nbx_readout_bwd", T, M * B, p.beta, _ptr(gy), _ptr(dZ), C.c_void_p(0), _stream())
        # dW_out = sum_t dZ_t S_t^T ; readout term of ds: R_t = W_out^T dZ_t
        dW_out = torch.matmul(dZ.permute(1, 0, 2).reshape(M, T * B), S.permute(1, 0, 2).reshape(N, T * B).t())
        W_out_t = W_out.t().contiguous()
        R = torch.matmul(W_out_t, dZ)  # [T, N, B]
        bits = _spike_words(N, B, x.device)
        dW_rec = torch.zeros_like(W_rec)
        dW_in = torch.zeros_like(W_in)
        dV = torch.zeros((N, B), device=x.device, dtype=torch.float32)
        dI_after = None
        zeros_s = torch.zeros((1, N, B), device=x.device, dtype=torch.float32)
        for seg in reversed(range(len(ckpts))):
            t0, t1 = seg * K, min(T, (seg + 1) * K)
            k = t1 - t0
            V = ckpts[seg].clone()
            Vm = torch.empty((k, N, B), device=x.device, dtype=torch.float32)
            S_scr = torch.empty((k, N, B), device=x.device, dtype=torch.float32)
            _fwd_segment(cp, IN, W_rec, W_rec_t, bits, S[t0 - 1] if t0 > 0 else zeros_s[0], V, t0, t1, S_scr,
                         Vm_out=Vm)
            dI = torch.empty((k, N, B), device=x.device, dtype=torch.float32)
            # steps t1-1 .. t0, each fusing G = W_rec^T dI_{t+1} with the adjoint step; dI_{t1} is
            # the first step of the later segment (none after the last step)
            L.call("nbx_lif_bwd_segment", C.byref(cp), _ptr(W_rec_t), _ptr(dI_after), _ptr(R[t0]), _ptr(Vm), _ptr(dV),
                   _ptr(dI), k, _stream())
            dI_after = dI[0]
            # spikes feeding each step of the segment: s_{t-1}
            S_prev = S[t0 - 1:t1 - 1] if t0 > 0 else torch.cat([zeros_s, S[0:t1 - 1]], dim=0)
            dI2 = dI.permute(1, 0, 2).reshape(N, k * B)
            dW_rec += torch.matmul(dI2, S_prev.permute(1, 0, 2).reshape(N, k * B).t())
            dW_in += torch.matmul(dI2, x[t0:t1].permute(1, 0, 2).reshape(cin, k * B).t())
        return dW_in, dW_rec, dW_out, None, None, None


def loss_and_grads(W_in, W_rec, W_out, x, target, p: LifParams, checkpoint: int = 50):
    """MSE loss of the readout against `target` and the gradients w.r.t. the three weight
    matrices (all torch tensors on the GPU). Returns (loss, [dW_in, dW_rec, dW_out], spikes)."""
    ws = [w.detach().clone().requires_grad_(True) for w in (W_in, W_rec, W_out)]
    y, S = LifAdjoint.apply(ws[0], ws[1], ws[2], x, p, checkpoint)
    loss = ((y - target) ** 2).mean()
    loss.backward()
    return loss.detach(), [w.grad for w in ws], S


class _Spike(torch.autograd.Function):
    """s = Vm > V_th (or the forced spikes), with the fast-sigmoid surrogate derivative."""

    @staticmethod
    def forward(ctx, vm, v_th: float, slope: float, forced):
        ctx.save_for_backward(vm)
        ctx.v_th, ctx.slope = v_th, slope
        return forced if forced is not None else (vm > v_th).to(vm.dtype)

    @staticmethod
    def backward(ctx, gs):
        (vm,) = ctx.saved_tensors
        den = 1.0 + ctx.slope * (vm - ctx.v_th).abs()
        return gs / (den * den), None, None, None


def autograd_loss_and_grads(W_in, W_rec, W_out, x, target, p: LifParams, spikes=None):
    """The plain path the adjoint replaces: the same network and loss unrolled step by step with
    torch operations (FP32 on the GPU, matmul through the library BLAS) and differentiated by
    torch autograd, holding every step's tensors. Surrogate mode only (p.mode == 1). `spikes`
    ([T, N, B]) forces the spike train, so gradients can be compared on the same trajectory.
    Returns (loss, [dW_in, dW_rec, dW_out], spikes)."""
    assert p.mode == 1, "the plain unroll implements the surrogate-gradient mode"
    ws = [w.detach().clone().requires_grad_(True) for w in (W_in, W_rec, W_out)]
    Wi, Wr, Wo = ws
    T, cin, B = x.shape
    N = Wr.shape[0]
    g = p.g
    IN = torch.matmul(Wi, x)  # [T, N, B]
    V = torch.full((N, B), p.v_rest, device=x.device, dtype=torch.float32)
    s = torch.zeros((N, B), device=x.device, dtype=torch.float32)
    y = None
    beta = p.beta
    loss = 0.0
    out = []
    for t in range(T):
        I = IN[t] + torch.matmul(Wr, s)
        vm = g * V + (1.0 - g) * (p.v_rest + p.r * I)
        s = _Spike.apply(vm, p.v_th, p.slope, spikes[t] if spikes is not None else None)
        V = vm - (p.v_th - p.v_reset) * s
        z = torch.matmul(Wo, s)
        y = (1.0 - beta) * z if y is None else beta * y + (1.0 - beta) * z
        loss = loss + ((y - target[t]) ** 2).sum()
        out.append(s.detach())
    loss = loss / target.numel()
    loss.backward()
    return loss.detach(), [w.grad for w in ws], torch.stack(out)
