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


@pytest.mark.parametrize("threads,exact", [(256, "0"), (1024, "0"), (1024, "1")],
                         ids=["256", "1024", "1024-scale-and-round"])
def test_batched_equals_per_instance_path(gpu, threads, exact, monkeypatch):
    # "1": every instance on the scale-and-round quantization instead of the single fma
    monkeypatch.setenv("NBX_BATCH_EXACT", exact)
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


def _assert_same_run(a, b, inst):
    for k, ins in enumerate(inst):
        assert np.array_equal(a.raster(k), b.raster(k)), f"instance {k} (n={ins.n}) raster differs"
        va, la = a.state(k)
        vb, lb = b.state(k)
        assert np.array_equal(va.view(np.uint32), vb.view(np.uint32)), f"instance {k} V differs"
        assert np.array_equal(la, lb)
    assert a.events() == b.events() > 0


@pytest.mark.parametrize("exact", ["0", "1"], ids=["fma", "scale-and-round"])
def test_split_schedule_equals_per_instance_path(gpu, exact, monkeypatch):
    # parts of 32 .. 960 columns (4 neurons per thread), with 1 .. 32 row groups sharing the fired rows
    monkeypatch.setenv("NBX_BATCH_EXACT", exact)
    T = 500
    inst, shared, b = make_batch(count=12, n_lo=300, n_hi=1500, seed=31)
    req = np.array([1, 2, 3, 5, 8, 13, 1, 4, 40, 2, 7, 3], np.int32)
    b.set_parts(req)
    got = b.parts()
    assert (got >= 1).all() and (got <= req).all() and got.max() > 1
    assert b.schedule()[0] == 256
    b.record(0, T)
    b.run(T)
    spikes = 0
    for k, ins in enumerate(inst):
        net = W.instance_workload(ins, shared).make_net(fixed_shift=ins.fixed_shift)
        net.record(0, T, counts=True, raster=True)
        net.run(T, L.ENGINE_GRAPH)
        assert np.array_equal(b.raster(k), net.raster()), f"instance {k} (n={ins.n}, {got[k]} parts) raster differs"
        v, last = b.state(k)
        x, lastp = net.get_state()
        assert np.array_equal(v.view(np.uint32), x[:, 0].view(np.uint32))
        assert np.array_equal(last, lastp)
        spikes += int(net.counts().sum())
        net.close()
    assert spikes > 300


def test_split_schedule_wide_instances(gpu):
    # instances wider than one part (2048 columns) must be split; switching schedules between
    # runs carries the last step's spikes over
    T = 300
    inst, shared = W.batched_lif(count=4, n_lo=2100, n_hi=5000, seed=17)
    a, b = Batch(inst, **shared), Batch(inst, **shared)
    a.set_schedule(1024, 64)
    a.record(0, T)
    a.run(T)
    with pytest.raises(L.NbxError) as e:
        b.set_parts([1] * len(inst))
    assert e.value.status == L.NBX_ERR_INVALID
    b.record(0, T)
    b.set_parts([3, 4, 6, 9])  # parts wider than 1024 columns: 8 neurons per thread
    b.run(70)
    b.set_schedule(1024, 50)
    b.run(110)
    b.set_parts([5, 5, 5, 12])  # all parts within 1024 columns: 4 neurons per thread
    b.run(T - 180)
    _assert_same_run(a, b, inst)


def test_autotuner_split_restores_state(gpu):
    T = 240
    inst, shared, a = make_batch(count=16, n_lo=200, n_hi=2600, seed=23)
    _, _, b = make_batch(count=16, n_lo=200, n_hi=2600, seed=23)
    a.tune(40)
    parts = a.parts()
    assert parts.shape == (16,)
    a.record(0, T)
    a.run(T)
    b.set_schedule(1024, 100)
    b.record(0, T)
    b.run(T)
    _assert_same_run(a, b, inst)
