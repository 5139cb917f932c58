"""GPU parity for the remaining legacy models (double- and triple-exponential conductance LIF,
PoissonRate, trigger relay) on every engine, in replay delivery, with dense signed weights and
from .nbnet tables. Bar: bit-exact against the oracle (state bits, LastSpike, rasters, ring)."""
import numpy as np
import pytest

from harness import OracleRun
from nbnet_file import records, write_nbnet
from neuronbridge_amd import _lib as L
from neuronbridge_amd import workloads as W
from test_gpu_parity import ENGINES, assert_bits_equal, compare, random_spikes, replay_expect

pytestmark = pytest.mark.gpu

CASES = [
    ("lif2-euler", L.LIF_EXP2, L.EULER),
    ("lif2-rk4", L.LIF_EXP2, L.RK4),
    ("lif3-euler", L.LIF_EXP3, L.EULER),
    ("lif3-rk4", L.LIF_EXP3, L.RK4),
    ("poisson", L.POISSON, L.EULER),
    ("relay", L.TRIGGER_RELAY, L.EULER),
]


@pytest.mark.parametrize("case,model,integrator", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("engine", list(ENGINES))
def test_model_closed_loop_bitexact(gpu, case, model, integrator, engine):
    wl = W.model_sparse(model, integrator=integrator)
    T = 300
    net = wl.make_net()
    net.record(0, T, counts=True, raster=True)
    net.run(T, ENGINES[engine])
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    counts_o, raster_o = run.run(T)
    assert 200 < counts_o.sum() < 0.5 * T * wl.spec.n, "workload too quiet or saturated to test delivery"
    compare(net, run, wl, counts_o, raster_o)


def test_three_channel_engines_interoperate(gpu):
    # NMDA arrivals (channel 2) carry over between the persistent engine's LDS ring and the device ring
    wl = W.model_sparse(L.LIF_EXP3, n=3000, k_in=120, integrator=L.RK4, seed=10)
    assert wl.spec.channels == 3
    net = wl.make_net()
    assert net.pick_engine() == L.ENGINE_PERSISTENT
    net.record(0, 260, counts=True, raster=True)
    net.run(90, L.ENGINE_PERSISTENT)
    net.run(70, L.ENGINE_GRAPH)
    net.run(100, L.ENGINE_PERSISTENT)
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    counts_o, raster_o = run.run(260)
    compare(net, run, wl, counts_o, raster_o)


@pytest.mark.parametrize("model,slots", [(L.LIF_EXP2, None), (L.LIF_EXP3, None), (L.POISSON, None),
                                         (L.LIF_EXP2, "1"), (L.LIF_EXP3, "1"), (L.LIF_EXP3, "0")],
                         ids=["lif2", "lif3", "poisson", "lif2-slots", "lif3-slots", "lif3-offsets"])
def test_model_replay_matches_brute_force(gpu, model, slots, monkeypatch):
    if slots is not None:  # force the fixed-slot or the offset-layout kernel
        monkeypatch.setenv("NBX_REPLAY_SLOTS", slots)
        monkeypatch.setenv("NBX_REPLAY_M", "1")
    wl = W.model_sparse(model, n=2500, k_in=80, seed=8)
    net = wl.make_net()
    rng = np.random.default_rng(11)
    K = 64
    spikes = random_spikes(rng, wl.spec.n, K, 0, 60)
    e0 = net.events()
    if slots is not None:
        assert net.replay_path() == (L.REPLAY_STREAM if slots == "1" else L.REPLAY_OFFSETS)
    out = net.replay(spikes)
    expect, events = replay_expect(wl, net.fixed_shift, spikes)
    assert out.shape[1] == wl.spec.channels
    assert_bits_equal(out.astype(np.int64), expect, "replay arrivals")
    assert net.events() - e0 == events


@pytest.mark.parametrize("engine", [L.ENGINE_GRAPH, L.ENGINE_PERSISTENT], ids=["graph", "persistent"])
@pytest.mark.parametrize("model", [L.LIF_EXP2, L.LIF_EXP3], ids=["lif2", "lif3"])
def test_dense_signed_weights(gpu, model, engine):
    # negative dense weights are inhibitory conductance inputs on their own channel
    n, T = 400, 400
    wl = W.model_sparse(model, n=n, k_in=10, seed=5)
    rng = np.random.default_rng(2)
    wt = rng.uniform(0.0, 0.004, (n, n)).astype(np.float32)
    wt[int(0.8 * n):] = -rng.uniform(0.0, 0.01, (n - int(0.8 * n), n)).astype(np.float32)
    wl.syn, wl.dense_wt, wl.dense_delay = None, wt, 2
    net = wl.make_net()
    net.record(0, T, counts=True, raster=True)
    net.run(T, engine)
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    counts_o, raster_o = run.run(T)
    assert counts_o.sum() > 100
    compare(net, run, wl, counts_o, raster_o)


def test_nbnet_nmda_synapses(gpu, tmp_path):
    wl = W.model_sparse(L.LIF_EXP3, n=1500, k_in=60, seed=12)
    n = wl.spec.n
    pre = np.repeat(np.arange(n, dtype=np.int64), np.diff(wl.syn.row_ptr))
    recs = records(pre, wl.syn.post, wl.syn.type, wl.syn.weight, wl.syn.delay)
    recs = recs[np.random.default_rng(1).permutation(len(recs))]
    path = str(tmp_path / "nmda.nbnet")
    write_nbnet(path, {"layers": []}, recs)
    assert (recs["type"] == 2).sum() > 1000
    net = wl.make_net(nbnet=path)
    T = 250
    net.record(0, T, counts=True, raster=True)
    net.run(T, L.ENGINE_GRAPH)
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    counts_o, raster_o = run.run(T)
    compare(net, run, wl, counts_o, raster_o)
    # a two-channel model has no NMDA input
    wl2 = W.model_sparse(L.LIF_EXP2, n=1500, k_in=60, seed=12)
    with pytest.raises(L.NbxError) as e:
        wl2.make_net(nbnet=path)
    assert e.value.status == L.NBX_ERR_UNSUPPORTED
