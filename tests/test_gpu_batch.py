"""Batched multi-instance engine: every instance's raster and final state equal the
per-instance FP32 path (single-network engine on the BF16-valued weights), bit for bit,
and the CPU oracle on sampled instances."""
import numpy as np
import pytest

from harness import OracleRun
from neuronbridge_amd import _lib as L
from neuronbridge_amd import workloads as W
from neuronbridge_amd.batch import Batch

pytestmark = pytest.mark.gpu


def make_batch(count=20, n_lo=64, n_hi=1500, seed=77):
    inst, shared = W.batched_lif(count=count, n_lo=n_lo, n_hi=n_hi, seed=seed)
    return inst, shared, Batch(inst, **shared)


@pytest.mark.parametrize("threads", [256, 1024])
def test_batched_equals_per_instance_path(gpu, threads):
    T = 600
    inst, shared, b = make_batch()
    b.set_schedule(threads, 128)
    b.record(0, T)
    b.run(T)
    total_spikes = 0
    for k, ins in enumerate(inst):
        wl = W.instance_workload(ins, shared)
        net = wl.make_net(fixed_shift=ins.fixed_shift)
        net.record(0, T, counts=True, raster=True)
        net.run(T, L.ENGINE_GRAPH)
        rb = b.raster(k)
        assert np.array_equal(rb, net.raster()), f"instance {k} (n={ins.n}) raster differs"
        v, last = b.state(k)
        x, lastp = net.get_state()
        assert np.array_equal(v.view(np.uint32), x[:, 0].view(np.uint32)), f"instance {k} V differs"
        assert np.array_equal(last, lastp)
        total_spikes += int(net.counts().sum())
        net.close()
    assert total_spikes > 500


def test_batched_against_oracle(gpu):
    T = 400
    inst, shared, b = make_batch(count=6, n_lo=200, n_hi=900, seed=5)
    b.record(0, T)
    b.run(T)
    for k in (0, 3):
        wl = W.instance_workload(inst[k], shared)
        run = OracleRun(wl, fixed_shift=inst[k].fixed_shift)
        _, raster = run.run(T)
        assert np.array_equal(b.raster(k), raster)
        v, last = b.state(k)
        assert np.array_equal(v.view(np.uint32), run.x[:, 0].view(np.uint32))


def test_autotuner_and_chunked_graph(gpu):
    # tuning restores the state; results do not depend on the schedule or chunking
    T = 300
    inst, shared, a = make_batch(count=10, seed=9)
    inst2, _, b = make_batch(count=10, seed=9)
    thr = a.tune(50)
    assert thr in (256, 512, 1024)
    a.record(0, T)
    a.run(T)
    b.set_schedule(512, 37)
    b.record(0, T)
    b.run(100)
    b.run(T - 100)
    for k in range(10):
        assert np.array_equal(a.raster(k), b.raster(k))
    assert a.events() == b.events() > 0
