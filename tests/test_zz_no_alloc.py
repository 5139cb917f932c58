"""The allocation-free replay-STDP entry point (nbx_replay_stdp_device_noalloc) against the
allocating one (nbx_replay_stdp_device), through the C ABI on device-resident rasters.

Two networks built from one workload replay the same raster: one through each entry point, in
two calls in a row, the second with a smaller input (fewer steps and spikes, chained with the
D - 1 lead steps). Weights, trace records and the update counters must be identical bit for bit
after each call, and the allocation-free calls must leave the network usable for the allocating
entry point afterwards (the static workspace keeps stale contents between calls).
"""
import ctypes as C

import numpy as np
import pytest
import torch  # noqa: F401  (torch's HIP runtime first, as in the other GPU tests)

from harness import raster_csr, raster_lists
from neuronbridge_amd import _lib as L
from neuronbridge_amd import workloads as W
from neuronbridge_amd.hiptime import DeviceBuffer

pytestmark = pytest.mark.gpu


def _state(net):
    return (net.weights(),) + net.stdp_state()


def _same(a, b, what):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    if a.dtype == np.float32:
        a, b = a.view(np.uint32), b.view(np.uint32)
    bad = np.nonzero(a != b)[0]
    assert bad.size == 0, f"{what}: {bad.size} mismatches, first at {bad[:3]}"


def _device_call(name, net, lists, t0, lead):
    ids, tp = raster_csr(lists)
    d_ids, d_tp, d_upd = DeviceBuffer(ids.nbytes), DeviceBuffer(tp.nbytes), DeviceBuffer(8)
    d_ids.upload(ids)
    d_tp.upload(tp)
    d_upd.upload(np.zeros(1, np.uint64))
    L.call(name, net.h, d_ids.ptr, d_tp.ptr, int(t0), int(lead), len(lists) - int(lead), int(tp[-1]), d_upd.ptr)
    net.sync()
    out = np.zeros(1, np.uint64)
    d_upd.download(out)
    return int(out[0])


def test_noalloc_replay_stdp_matches_allocating_entry_point(gpu):
    T1, T2 = 700, 150
    wl = W.stdp_izh(n=1500, k_in=60, steps=T1 + T2, w_bg=2.0)
    src = wl.make_net()
    src.record(0, T1 + T2, counts=True, raster=True)
    src.run(T1 + T2, L.ENGINE_GRAPH)
    lists = raster_lists(src.raster(), wl.spec.n)
    src.close()
    assert sum(len(x) for x in lists) > 200
    D = wl.d_slots
    a, b = wl.make_net(), wl.make_net()
    # first call: T1 steps
    ua = _device_call("nbx_replay_stdp_device", a, lists[:T1], 0, 0)
    ub = _device_call("nbx_replay_stdp_device_noalloc", b, lists[:T1], 0, 0)
    assert ua == ub > 0
    for x, y, what in zip(_state(a), _state(b), ("weights", "Apre", "Apost", "LastUpdate")):
        _same(x, y, "first call: " + what)
    # second call, smaller: T2 steps chained after the first with D - 1 lead steps
    lead = D - 1
    chunk = lists[T1 - lead:T1 + T2]
    ua = _device_call("nbx_replay_stdp_device", a, chunk, T1, lead)
    ub = _device_call("nbx_replay_stdp_device_noalloc", b, chunk, T1, lead)
    assert ua == ub > 0
    for x, y, what in zip(_state(a), _state(b), ("weights", "Apre", "Apost", "LastUpdate")):
        _same(x, y, "second call: " + what)
    # the network that used the static workspace still runs the allocating path identically
    tail = lists[T1 + T2 - 100 - lead:T1 + T2]
    ua = _device_call("nbx_replay_stdp_device", a, tail, T1 + T2 - 100, lead)
    ub = _device_call("nbx_replay_stdp_device", b, tail, T1 + T2 - 100, lead)
    assert ua == ub
    for x, y, what in zip(_state(a), _state(b), ("weights", "Apre", "Apost", "LastUpdate")):
        _same(x, y, "after: " + what)
    a.close()
    b.close()


def test_noalloc_replay_stdp_rejects_oversized_workspace(gpu):
    # 12 bytes per neuron and 64-step block: 20000 neurons x 1024 blocks (65535 steps) needs ~250 MB
    wl = W.stdp_izh(n=20000, k_in=20, steps=10, w_bg=2.0)
    net = wl.make_net()
    K = 65535
    tp = np.zeros(K + 1, np.int32)
    d_ids, d_tp = DeviceBuffer(4), DeviceBuffer(tp.nbytes)
    d_tp.upload(tp)
    with pytest.raises(L.NbxError, match="workspace"):
        L.call("nbx_replay_stdp_device_noalloc", net.h, d_ids.ptr, d_tp.ptr, 0, 0, K, 0, None)
    net.close()
