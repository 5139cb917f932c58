"""Surrogate-gradient adjoint on the GPU against the FP64 checker (autograd unroll with the
GPU's spike raster replayed, and finite differences in the smooth mode). Bar: gradients
within 1e-4 relative of FP64."""
import numpy as np
import pytest
import torch

from oracle import adjoint_ref as R
from neuronbridge_amd import workloads as W
from neuronbridge_amd.adjoint import LifParams, loss_and_grads

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def to_dev(*arrs):
    return [torch.tensor(a, device=DEV) for a in arrs]


def rel_err(a: torch.Tensor, b: torch.Tensor) -> float:
    a, b = a.double().cpu(), b.double().cpu()
    return float((a - b).norm() / b.norm())


def check_against_fp64(n, b, T, checkpoint, mode=1, seed=None, tol=1e-4):
    kw = dict(n=n, batch=b, steps=T)
    if seed is not None:
        kw["seed"] = seed
    Wi, Wr, Wo, x, tg = W.diff_lif(**kw)
    p = LifParams(mode=mode)
    loss, grads, S = loss_and_grads(*to_dev(Wi, Wr, Wo, x, tg), p, checkpoint)
    torch.cuda.synchronize()
    raster = S.cpu().double()
    # FP64 forward without replay: how many threshold decisions differ from the FP32 GPU run
    _, r64, _ = R.run(*[torch.tensor(a, dtype=torch.float64) for a in (Wi, Wr, Wo, x)], p)
    flips = int((r64 != raster).sum()) if mode != 2 else 0
    g64, loss64, _ = R.gradients(torch.tensor(Wi), torch.tensor(Wr), torch.tensor(Wo), torch.tensor(x), p,
                                 torch.tensor(tg), spikes=raster if mode != 2 else None)
    errs = [rel_err(g, r) for g, r in zip(grads, g64)]
    assert abs(float(loss) - loss64) <= 1e-5 * abs(loss64), (float(loss), loss64)
    assert max(errs) <= tol, errs
    return errs, flips, float(raster.mean())


@pytest.mark.parametrize("checkpoint", [0, 50])
def test_surrogate_gradients_match_fp64(gpu, checkpoint):
    errs, flips, rate = check_against_fp64(512, 16, 150, checkpoint)
    assert rate > 1e-3
    print(f"relative gradient errors {errs}, FP32/FP64 spike decisions differing {flips}")


def test_smooth_mode_matches_fp64_and_finite_differences(gpu):
    n, b, T = 96, 4, 60
    Wi, Wr, Wo, x, tg = W.diff_lif(n=n, batch=b, steps=T, seed=123)
    p = LifParams(mode=2)
    _, grads, _ = loss_and_grads(*to_dev(Wi, Wr, Wo, x, tg), p, 16)
    g64, _, _ = R.gradients(torch.tensor(Wi), torch.tensor(Wr), torch.tensor(Wo), torch.tensor(x), p, torch.tensor(tg))
    for which in range(3):
        assert rel_err(grads[which], g64[which]) <= 1e-4
        flat = grads[which].reshape(-1).double().cpu()
        idx = np.random.default_rng(which).choice(flat.numel(), size=6, replace=False)
        fd, _ = R.finite_difference(torch.tensor(Wi), torch.tensor(Wr), torch.tensor(Wo), torch.tensor(x), p,
                                    torch.tensor(tg), which, idx, eps=1e-6)
        assert torch.max(torch.abs(flat[idx] - fd)) <= 1e-4 * torch.max(torch.abs(fd)) + 1e-8


def test_checkpointing_does_not_change_gradients(gpu):
    Wi, Wr, Wo, x, tg = W.diff_lif(n=256, batch=8, steps=120, seed=9)
    p = LifParams()
    ref = None
    for k in (0, 50, 7):
        loss, grads, _ = loss_and_grads(*to_dev(Wi, Wr, Wo, x, tg), p, k)
        if ref is None:
            ref = grads
        else:
            for g, r in zip(grads, ref):
                assert torch.allclose(g, r, rtol=1e-6, atol=1e-9)


def test_full_size_configuration(gpu):
    # N=2048, 64 trials, T=500, checkpoint every 50 steps (the configuration's sizes)
    errs, flips, rate = check_against_fp64(2048, 64, 500, 50)
    print(f"full size: relative gradient errors {errs}, spike decisions differing {flips}, rate {rate * 1000:.1f} Hz")


@pytest.mark.parametrize("n,b", [(2048, 64), (100, 5), (96, 4), (300, 33), (2050, 37), (1030, 64)])
def test_fused_recurrent_steps_against_fp64(gpu, n, b):
    # the fused step kernels (W s on FP32 matrix cores + the step) against the same step in FP64,
    # at the configuration's shape and at shapes that exercise partial tiles, a partial last
    # reduction chunk and unaligned rows
    import ctypes as C
    from neuronbridge_amd import _lib as L
    rng = np.random.default_rng(n + b)
    p = LifParams()
    cp = p.c_params(n, b)
    W = rng.standard_normal((n, n)) / np.sqrt(n)
    s_prev = (rng.random((n, b)) < 0.05).astype(np.float64)
    IN = rng.standard_normal((n, b)) * 0.05
    V0 = rng.uniform(-0.2, 0.9, (n, b))
    Wd, sd, INd, Vd = to_dev(W.astype(np.float32), s_prev.astype(np.float32), IN.astype(np.float32),
                             V0.astype(np.float32))
    S = torch.empty((1, n, b), device=DEV)
    Vm = torch.empty((1, n, b), device=DEV)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    ptr = lambda t: C.c_void_p(t.data_ptr())
    L.call("nbx_lif_fwd_segment", C.byref(cp), ptr(Wd), ptr(INd), ptr(sd), ptr(Vd), ptr(S), ptr(Vm), 1, stream)
    torch.cuda.synchronize()
    Wf, sf, INf, V0f = (a.astype(np.float32).astype(np.float64) for a in (W, s_prev, IN, V0))
    vm64 = p.g * V0f + (1 - p.g) * (p.v_rest + p.r * (INf + Wf @ sf))
    vm = Vm[0].double().cpu().numpy()
    assert np.max(np.abs(vm - vm64)) <= 1e-5 * max(1.0, np.max(np.abs(vm64)))
    near = np.abs(vm64 - p.v_th) < 1e-4  # threshold decisions away from rounding distance agree
    assert np.array_equal((S[0].cpu().numpy() > 0)[~near], (vm64 > p.v_th)[~near])
    # one adjoint step: dI = (1-g) r (dV + sigma'(Vm - V_th) (R + W^T dI_next - (V_th - V_reset) dV))
    dI_next = rng.standard_normal((n, b)) * 0.1
    Rr = rng.standard_normal((n, b)) * 0.1
    dV = rng.standard_normal((n, b)) * 0.1
    WTd, dInd, Rd, dVd = to_dev(np.ascontiguousarray(W.T).astype(np.float32), dI_next.astype(np.float32),
                                Rr.astype(np.float32), dV.astype(np.float32))
    dI = torch.empty((1, n, b), device=DEV)
    L.call("nbx_lif_bwd_segment", C.byref(cp), ptr(WTd), ptr(dInd), ptr(Rd), ptr(Vm), ptr(dVd), ptr(dI), 1, stream)
    torch.cuda.synchronize()
    dIn_f, R_f, dV_f = (a.astype(np.float32).astype(np.float64) for a in (dI_next, Rr, dV))
    ds = R_f + Wf.T @ dIn_f - (p.v_th - p.v_reset) * dV_f
    sg = 1.0 / (1.0 + p.slope * np.abs(vm - p.v_th)) ** 2
    dvm = dV_f + sg * ds
    want = (1 - p.g) * p.r * dvm
    got = dI[0].double().cpu().numpy()
    assert np.max(np.abs(got - want)) <= 1e-5 * max(1.0, np.max(np.abs(want)))
    assert np.max(np.abs(dVd.double().cpu().numpy() - p.g * dvm)) <= 1e-5 * max(1.0, np.max(np.abs(dvm)))


@pytest.mark.parametrize("n,b", [(2048, 64), (100, 5), (300, 33), (2100, 9)])
def test_sparse_forward_segment_against_fp64_and_dense(gpu, n, b):
    # the forward segment for binary spikes (sum of the W_rec^T rows of the neurons that fired)
    # against the same steps in FP64 and against the dense MFMA segment, over several steps so the
    # spike words of one step feed the next; recomputation is bit-identical
    import ctypes as C
    from neuronbridge_amd import _lib as L
    rng = np.random.default_rng(n * 7 + b)
    p = LifParams()
    cp = p.c_params(n, b)
    k = 6
    W = (rng.standard_normal((n, n)) / np.sqrt(n)).astype(np.float32)
    s0 = (rng.random((n, b)) < 0.05).astype(np.float32)
    IN = (rng.standard_normal((k, n, b)) * 0.08 + 0.04).astype(np.float32)
    V0 = rng.uniform(-0.2, 0.95, (n, b)).astype(np.float32)
    Wd, WTd, s0d, INd = to_dev(W, np.ascontiguousarray(W.T), s0, IN)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    ptr = lambda t: C.c_void_p(t.data_ptr())
    bits = torch.empty(2 * b * ((n + 31) // 32), device=DEV, dtype=torch.int32)
    outs = []
    for rep in range(2):
        V = torch.tensor(V0, device=DEV)
        S = torch.empty((k, n, b), device=DEV)
        Vm = torch.empty((k, n, b), device=DEV)
        L.call("nbx_lif_fwd_segment_sparse", C.byref(cp), ptr(WTd), ptr(INd), ptr(s0d), ptr(V), ptr(S), ptr(Vm),
               ptr(bits), k, stream)
        torch.cuda.synchronize()
        outs.append((V.cpu().numpy(), S.cpu().numpy(), Vm.cpu().numpy()))
    for x, y in zip(outs[0], outs[1]):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), "recomputed segment differs"
    V_g, S_g, Vm_g = outs[0]
    # FP64 steps fed with the GPU's own spikes (threshold decisions checked away from rounding distance)
    Wf = W.astype(np.float64)
    v = V0.astype(np.float64)
    s_prev = s0.astype(np.float64)
    for t in range(k):
        vm64 = p.g * v + (1 - p.g) * (p.v_rest + p.r * (IN[t].astype(np.float64) + Wf @ s_prev))
        assert np.max(np.abs(Vm_g[t] - vm64)) <= 1e-5 * max(1.0, np.max(np.abs(vm64)))
        near = np.abs(vm64 - p.v_th) < 1e-4
        assert np.array_equal((S_g[t] > 0)[~near], (vm64 > p.v_th)[~near])
        s_prev = S_g[t].astype(np.float64)
        v = vm64 - (p.v_th - p.v_reset) * s_prev
    assert S_g.sum() > 0
    # the dense segment on the same inputs gives the same spikes away from threshold and the same
    # potentials to rounding
    V = torch.tensor(V0, device=DEV)
    S = torch.empty((k, n, b), device=DEV)
    Vm = torch.empty((k, n, b), device=DEV)
    L.call("nbx_lif_fwd_segment", C.byref(cp), ptr(Wd), ptr(INd), ptr(s0d), ptr(V), ptr(S), ptr(Vm), k, stream)
    torch.cuda.synchronize()
    assert np.max(np.abs(Vm.cpu().numpy()[0] - Vm_g[0])) <= 1e-5
    # mode 2 is not binary: rejected
    cp2 = LifParams(mode=2).c_params(n, b)
    with pytest.raises(L.NbxError):
        L.call("nbx_lif_fwd_segment_sparse", C.byref(cp2), ptr(WTd), ptr(INd), ptr(s0d), ptr(V), ptr(S), ptr(Vm),
               ptr(bits), k, stream)
