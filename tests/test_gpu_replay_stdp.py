"""Replay STDP (nbx_replay_stdp) on the GPU: given the raster of a closed-loop run, the one-pass
kernel leaves the same weights and trace records, bit for bit, as the closed loop did, and as the
oracle's replay (orc_replay_stdp) does.

The rasters come from closed-loop runs of the STDP network itself (GPU engine), so the test also
covers networks whose spike schedule a test could not plan by hand.
"""
import ctypes as C

import numpy as np
import pytest
import torch  # before libnbx starts HIP: torch brings its own HIP runtime, which must load first

from harness import OracleRun, raster_csr, raster_lists
from neuronbridge_amd import _lib as L
from neuronbridge_amd import workloads as W

pytestmark = pytest.mark.gpu


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def assert_same(a, b, what):
    bad = np.nonzero(bits(a) != bits(b))[0]
    assert bad.size == 0, f"{what}: {bad.size} mismatches, first at {bad[:3]}: {a[bad[:3]]} vs {b[bad[:3]]}"


def closed_loop(wl, T, engine=L.ENGINE_GRAPH):
    net = wl.make_net()
    net.record(0, T, counts=True, raster=True)
    net.run(T, engine)
    out = (net.weights(),) + net.stdp_state()
    lists = raster_lists(net.raster(), wl.spec.n)
    net.close()
    return out, lists


def replayed_state(net):
    return (net.weights(),) + net.stdp_state()


@pytest.mark.parametrize("engine", ["kernels", "graph"])
def test_replay_stdp_equals_closed_loop_and_oracle(gpu, engine):
    T = 1200
    wl = W.stdp_izh(n=1500, k_in=60, steps=T, w_bg=2.0)
    ref, lists = closed_loop(wl, T, L.ENGINE_KERNELS if engine == "kernels" else L.ENGINE_GRAPH)
    assert sum(len(x) for x in lists) > 200
    net = wl.make_net()
    upd = net.replay_stdp(lists)
    got = replayed_state(net)
    for a, b, name in zip(got, ref, ("weights", "Apre", "Apost", "LastUpdate")):
        assert_same(a, b, name)
    assert not np.array_equal(got[0], wl.syn.weight)
    run = OracleRun(wl)
    assert run.replay_stdp(lists) == upd
    assert_same(got[0], run.w, "weights vs oracle replay")
    assert_same(got[3], run.slast, "LastUpdate vs oracle replay")
    # the replay touches neither the step counter nor non-plastic synapses
    assert net.time == 0
    keep = wl.syn.plastic == 0
    assert np.array_equal(got[0][keep], wl.syn.weight[keep])


def test_replay_stdp_chained_calls(gpu):
    T = 900
    wl = W.stdp_izh(n=1500, k_in=60, steps=T, w_bg=2.0)
    ref, lists = closed_loop(wl, T)
    D = wl.d_slots
    net = wl.make_net()
    total = 0
    t = 0
    for k in (1, 300, 37, T - 338):  # lead = D - 1 steps re-read for their arrivals
        lead = min(t, D - 1)
        total += net.replay_stdp(lists[t - lead:t + k], t0=t, lead=lead)
        t += k
    got = replayed_state(net)
    for a, b, name in zip(got, ref, ("weights", "Apre", "Apost", "LastUpdate")):
        assert_same(a, b, name)
    fresh = wl.make_net()
    assert fresh.replay_stdp(lists) == total


def test_replay_stdp_full_size_network(gpu):
    # the STDP configuration's network: 50,000 Izhikevich neurons, in-degree 500, delays 1..10
    T = 400
    wl = W.stdp_izh(steps=T)
    ref, lists = closed_loop(wl, T)
    net = wl.make_net()
    upd = net.replay_stdp(lists)
    assert upd > 1_000_000
    got = replayed_state(net)
    for a, b, name in zip(got, ref, ("weights", "Apre", "Apost", "LastUpdate")):
        assert_same(a, b, name)


def test_replay_stdp_edge_cases(gpu):
    wl = W.stdp_izh(n=600, k_in=40, steps=10, w_bg=2.0)
    n = wl.spec.n
    net = wl.make_net()
    w0 = net.weights()
    assert net.replay_stdp([[] for _ in range(50)]) == 0  # empty raster: nothing changes
    assert np.array_equal(net.weights(), w0)
    # spikes only on the last step: post-side updates only (the arrivals fall after the window)
    last = [[] for _ in range(49)] + [list(range(0, n, 7))]
    upd = net.replay_stdp(last, t0=50)
    run = OracleRun(wl)
    run.replay_stdp([[] for _ in range(50)])
    assert run.replay_stdp(last, t0=50) == upd > 0
    assert_same(net.weights(), run.w, "weights")
    assert_same(net.stdp_state()[2], run.slast, "LastUpdate")
    w1 = net.weights()
    with pytest.raises(RuntimeError, match="twice in one step"):
        net.replay_stdp([[3, 5, 3]], t0=200)
    with pytest.raises(RuntimeError, match="out of range"):
        net.replay_stdp([[n]], t0=200)
    with pytest.raises(RuntimeError, match="bad replay arguments"):
        net.replay_stdp([[1], [2]], t0=0, lead=1)  # lead steps before step 0
    with pytest.raises(RuntimeError, match="65535"):
        net.replay_stdp([[]] * 65536, t0=300)  # local steps are 16-bit: longer rasters are chained
    assert np.array_equal(net.weights(), w1)  # rejected calls change nothing
    plain = W.sparse_lif(n=500, k_in=20, steps=10).make_net()
    with pytest.raises(RuntimeError, match="STDP"):
        plain.replay_stdp([[1]])


def test_replay_stdp_device_entry_rejects_bad_raster(gpu):
    wl = W.stdp_izh(n=600, k_in=40, steps=10, w_bg=2.0)
    net = wl.make_net()
    w0 = net.weights()
    ids, tp = raster_csr([[4, 9, 4], [1]])  # neuron 4 twice in step 0
    d_ids = torch.from_numpy(ids).cuda()
    d_tp = torch.from_numpy(tp).cuda()
    torch.cuda.synchronize()
    L.call("nbx_replay_stdp_device", net.h, C.c_void_p(d_ids.data_ptr()), C.c_void_p(d_tp.data_ptr()), 0, 0, 2,
           int(tp[-1]), None)
    with pytest.raises(RuntimeError, match="invalid raster"):
        net.sync()
    assert np.array_equal(net.weights(), w0)
    net.sync()  # reported once
    # a valid raster through the same entry point, with its update counter
    ids, tp = raster_csr([[4, 9], [1, 4], [], [2]])
    d_ids = torch.from_numpy(ids).cuda()
    d_tp = torch.from_numpy(tp).cuda()
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    L.call("nbx_replay_stdp_device", net.h, C.c_void_p(d_ids.data_ptr()), C.c_void_p(d_tp.data_ptr()), 0, 0, 4,
           int(tp[-1]), C.c_void_p(cnt.data_ptr()))
    net.sync()
    run = OracleRun(wl)
    assert run.replay_stdp([[4, 9], [1, 4], [], [2]]) == int(cnt.item()) > 0
    assert_same(net.weights(), run.w, "weights")


@pytest.mark.parametrize("p_spike", [0.002, 0.03, 0.2], ids=["sparse", "medium", "dense"])
def test_replay_stdp_random_rasters_against_oracle(gpu, p_spike):
    """Synthetic rasters of every density: short lists (merged from LDS), lists of 20 and more
    spikes and rows of 64 and more (merged from memory), against orc_replay_stdp, in two chained
    calls."""
    T = 400
    wl = W.stdp_izh(n=1200, k_in=50, steps=10, w_bg=2.0)
    n, D = wl.spec.n, wl.d_slots
    rng = np.random.default_rng(int(p_spike * 1e4))
    lists = [np.nonzero(rng.random(n) < p_spike)[0].astype(np.int32) for _ in range(T)]
    net = wl.make_net()
    run = OracleRun(wl)
    t1 = 173
    for t0, lead, part in ((0, 0, lists[:t1]), (t1, D - 1, lists[t1 - (D - 1):])):
        assert net.replay_stdp(part, t0=t0, lead=lead) == run.replay_stdp(part, t0=t0, lead=lead)
    assert_same(net.weights(), run.w, "weights")
    apre, apost, last = net.stdp_state()
    assert_same(apre, run.apre, "Apre")
    assert_same(apost, run.apost, "Apost")
    assert_same(last, run.slast, "LastUpdate")


def test_replay_stdp_long_decay_tables(gpu):
    # dt 0.1 ms: the decay tables (exp(-k dt / 20 ms) down to 0) are longer than the part staged in LDS
    T = 600
    wl = W.stdp_izh(n=1500, k_in=60, steps=T, w_bg=2.0)
    wl.spec.dt = 0.1
    ref, lists = closed_loop(wl, T)
    assert sum(len(x) for x in lists) > 100
    net = wl.make_net()
    net.replay_stdp(lists[:300])
    net.replay_stdp(lists[300 - (wl.d_slots - 1):], t0=300, lead=wl.d_slots - 1)
    got = replayed_state(net)
    for a, b, name in zip(got, ref, ("weights", "Apre", "Apost", "LastUpdate")):
        assert_same(a, b, name)
