"""The benchmark configurations at their full sizes against the CPU oracle, bit for bit.

  sparse LIF   1e5 neurons, in-degree 1000 (1e8 synapses), delays U{1..20}, RK4 dt 0.1 ms:
               300 closed-loop steps on the persistent and the graph engine, state, LastSpike,
               spike counts, raster, delay ring and event count equal to the oracle's
  STDP         5e4 Izhikevich neurons, in-degree 500 (2.5e7 synapses, 1.6e7 plastic), delays
               U{1..10}, dt 0.5 ms: 2000 closed-loop steps; final weights, STDP traces, state and
               raster equal to the oracle's
  batched      256 LIF instances, N ~ U{512..8192}, dense BF16 weights, RK4 dt 0.1 ms: 500 steps of
               the tuned batch schedule; the rasters and final state of 8 instances spread over
               the size range (the largest included) equal the plain per-instance FP32 path, and
               two of them the oracle
"""
import numpy as np
import pytest

from harness import OracleRun
from neuronbridge_amd import _lib as L
from neuronbridge_amd import workloads as W
from neuronbridge_amd.batch import Batch
from test_gpu_parity import assert_bits_equal, compare

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2():
    wl = W.sparse_lif(steps=300)
    T = 300
    run = OracleRun(wl, fixed_shift=23)
    counts, raster = run.run(T)
    return wl, run, counts, raster


@pytest.mark.parametrize("engine", [L.ENGINE_PERSISTENT, L.ENGINE_GRAPH], ids=["persistent", "graph"])
def test_sparse_lif_full_size_vs_oracle(gpu, c2, engine):
    wl, run, counts, raster = c2
    assert wl.spec.n == 100_000 and wl.syn.count == 100_000_000
    net = wl.make_net()
    assert net.fixed_shift == run.shift
    net.record(0, 300, counts=True, raster=True)
    net.run(300, engine)
    assert counts.sum() > 10_000
    compare(net, run, wl, counts, raster)
    net.close()


def test_stdp_full_size_vs_oracle(gpu):
    T = 2000
    wl = W.stdp_izh(steps=T)
    assert wl.spec.n == 50_000 and wl.syn.count == 25_000_000
    net = wl.make_net()
    net.record(0, T, counts=True, raster=True)
    net.run(T, L.ENGINE_GRAPH)
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    counts, raster = run.run(T)
    assert counts.sum() > 100_000
    compare(net, run, wl, counts, raster, ring=False)
    w = net.weights()
    assert_bits_equal(w, run.w, "weights after 2000 steps")
    apre, apost, last = net.stdp_state()
    assert_bits_equal(apre, run.apre, "Apre")
    assert_bits_equal(apost, run.apost, "Apost")
    assert_bits_equal(last, run.slast, "LastUpdate")
    plastic = wl.syn.plastic != 0
    assert plastic.sum() > 15_000_000
    assert (w[plastic] != wl.syn.weight[plastic]).mean() > 0.5, "STDP changed too few weights"
    net.close()


def test_batched_full_size_vs_plain_and_oracle(gpu):
    T = 500
    inst, shared = W.batched_lif(steps=T)
    assert len(inst) == 256
    b = Batch(inst, **shared)
    b.tune(50)
    b.record(0, T)
    b.run(T)
    order = np.argsort([i.n for i in inst])
    sample = order[np.linspace(0, len(inst) - 1, 8).astype(int)]
    assert inst[int(sample[-1])].n == max(i.n for i in inst)
    spikes = 0
    for k in sample:
        ins = inst[int(k)]
        net = W.instance_workload(ins, shared).make_net(fixed_shift=ins.fixed_shift)
        net.record(0, T, counts=True, raster=True)
        net.run(T, L.ENGINE_GRAPH)
        assert np.array_equal(b.raster(int(k)), net.raster()), f"instance {k} (n={ins.n}) raster differs"
        v, last = b.state(int(k))
        x, lastp = net.get_state()
        assert_bits_equal(v, x[:, 0], f"instance {k} V")
        assert_bits_equal(last, lastp, f"instance {k} LastSpike")
        spikes += int(net.counts().sum())
        net.close()
    assert spikes > 1000
    for k in (sample[1], sample[-1]):  # the oracle on a small and the largest instance
        ins = inst[int(k)]
        run = OracleRun(W.instance_workload(ins, shared), fixed_shift=ins.fixed_shift)
        _, raster = run.run(T)
        assert np.array_equal(b.raster(int(k)), raster), f"instance {k} (n={ins.n}) differs from the oracle"
        v, _ = b.state(int(k))
        assert_bits_equal(v, run.x[:, 0], f"instance {k} V vs oracle")


def test_fixed_point_steps_against_float_event_delivery_full_size(gpu):
    # the sparse LIF network at full size: at several points of a GPU run the oracle takes the GPU's
    # state and firing history and runs one step with the reference's per-event float delivery;
    # the GPU's own step (int32 fixed-point arrivals) agrees to 1e-5 relative (threshold decisions
    # within rounding distance of V_th may differ and are counted, not compared)
    import oracle as O
    wl = W.sparse_lif(steps=400)
    net = wl.make_net()
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    D, n = wl.d_slots, wl.spec.n
    t, worst, flips, compared = 0, 0.0, 0, 0
    for target in (5, 40, 120, 250, 390):
        net.run(target - t, L.ENGINE_PERSISTENT)
        t = target
        x, last = net.get_state()
        run.x[:] = x
        run.last[:] = last
        run.t = t
        run.hist[:] = 0
        run.hist_cnt[:] = 0
        for s in range(max(0, t - D + 1), t):
            ids = np.sort(net.history(s))
            run.hist[s % D, :ids.size] = ids
            run.hist_cnt[s % D] = ids.size
        _, ras = run.run(1, O.ACC_FLOAT_EVENTS, ring=False, raster=True)
        net.record(t, 1, counts=True, raster=True)
        net.run(1, L.ENGINE_GRAPH)
        t += 1
        xg, lastg = net.get_state()
        fo = np.unpackbits(ras[0].view(np.uint8), bitorder="little")[:n].astype(bool)
        fg = np.unpackbits(net.raster()[0].view(np.uint8), bitorder="little")[:n].astype(bool)
        same = fo == fg
        flips += int((~same).sum())
        rel = np.abs(xg[same, 0] - run.x[same, 0]) / np.maximum(np.abs(run.x[same, 0]), 1.0)
        worst = max(worst, float(rel.max()))
        compared += int(same.sum())
        assert fg.sum() > 10
        assert np.array_equal(lastg[same], run.last[same])
    print(f"fixed point vs float events at full size: worst relative step difference {worst:.3g}, "
          f"{flips} threshold decisions of {5 * n} differ")
    assert worst <= 1e-5
    assert flips <= 5
    net.close()
