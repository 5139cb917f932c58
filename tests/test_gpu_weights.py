"""Weight uploads and injected spikes between steps, against the CPU oracle.

A weight change reaches every arrival not yet delivered (the reference reads a synapse's weight
when the arrival is processed, PropogatedSpikeGroup.cpp:33-71): the oracle delivers at arrival
with the weights current then, the device moves its pending ring entries to the new weights.
"""
import numpy as np
import pytest

from harness import OracleRun
from neuronbridge_amd import _lib as L
from test_gpu_parity import ENGINES, assert_bits_equal, small_sparse

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("engine", ["kernels", "graph", "persistent"])
def test_set_weights_mid_run_bitexact(gpu, engine):
    wl = small_sparse(L.LIF_VJ, L.RK4, n=2000, k_in=100, seed=31)
    T1, T2 = 150, 150
    net = wl.make_net()
    net.record(0, T1 + T2, counts=True, raster=True)
    net.run(T1, ENGINES[engine])
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    c1, r1 = run.run(T1)
    rng = np.random.default_rng(5)
    w_new = (wl.syn.weight * rng.uniform(0.5, 1.6, wl.syn.count)).astype(np.float32)
    net.set_weights(w_new)
    run.w[:] = w_new  # the oracle delivers with the weights current at arrival
    j = int(wl.syn.row_ptr[7])
    net.set_synapse_weight(j, 0.25)
    run.w[j] = np.float32(0.25)
    assert net.synapse_weight(j) == np.float32(0.25)
    net.run(T2, ENGINES[engine])
    c2, r2 = run.run(T2, ring=False)
    assert c1.sum() + c2.sum() > 200
    x, last = net.get_state()
    assert_bits_equal(x, run.x, "state")
    assert_bits_equal(last, run.last, "LastSpike")
    assert_bits_equal(net.counts(), np.concatenate([c1, c2]), "counts")
    assert_bits_equal(net.raster(), np.concatenate([r1, r2]), "raster")
    assert_bits_equal(net.weights(), run.w, "weights")


def test_set_weights_rejections(gpu):
    wl = small_sparse(L.LIF_VJ, L.EULER, n=500, k_in=50, seed=3)
    net = wl.make_net()
    net.run(10, L.ENGINE_GRAPH)
    w0 = net.weights()
    bad = w0.copy()
    bad[3] = np.nan
    with pytest.raises(L.NbxError, match="finite"):
        net.set_weights(bad)
    huge = np.full_like(w0, 1e6)  # sum of |w| * 2^shift far beyond int32
    with pytest.raises(L.NbxError, match="overflow"):
        net.set_weights(huge)
    with pytest.raises(L.NbxError, match="overflow"):
        net.set_synapse_weight(0, 1e9)
    assert_bits_equal(net.weights(), w0, "weights unchanged after rejected uploads")
    net.begin_step(net.time)
    with pytest.raises(L.NbxError):
        net.set_weights(w0)
    net.run_propagation()
    net.neuron_step()
    net.commit_firings()


def test_fixed_shift_request_checked_at_finalize(gpu):
    wl = small_sparse(L.LIF_VJ, L.EULER, n=500, k_in=50, seed=3)
    with pytest.raises(L.NbxError, match="overflow"):
        wl.make_net(fixed_shift=30)


def test_inject_rejects_duplicates(gpu):
    wl = small_sparse(L.LIF_VJ, L.EULER, n=500, k_in=50, seed=3)
    net = wl.make_net()
    for _ in range(40):
        t = net.time
        net.begin_step(t)
        net.run_propagation()
        fired = net.neuron_step()
        if fired.size:
            with pytest.raises(L.NbxError, match="at most once"):
                net.inject_spikes([int(fired[0])])
        quiet = [i for i in range(20) if i not in set(fired.tolist())][:2]
        with pytest.raises(L.NbxError, match="at most once"):
            net.inject_spikes([quiet[0], quiet[0]])
        net.inject_spikes(quiet)
        with pytest.raises(L.NbxError, match="at most once"):
            net.inject_spikes(quiet[:1])
        net.commit_firings()
        assert sorted(net.history(t).tolist()) == sorted(set(fired.tolist()) | set(quiet))
