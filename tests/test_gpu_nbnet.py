"""Networks loaded from .nbnet connection tables (device-sorted) against the oracle."""
import numpy as np
import pytest

from harness import OracleRun
from nbnet_file import records, write_nbnet
from neuronbridge_amd import _lib as L
from neuronbridge_amd import workloads as W
from test_gpu_parity import assert_bits_equal, compare, small_sparse

pytestmark = pytest.mark.gpu

META = {"layers": [], "learning_rules": [], "outer_dynamics": [], "outer_dynamic_connections": [], "input_convs": []}


def file_from_csr(path, syn, seed=0, rule=None):
    """The workload's synapses as .nbnet records in a shuffled order."""
    n = syn.row_ptr.shape[0] - 1
    pre = np.repeat(np.arange(n, dtype=np.int64), np.diff(syn.row_ptr))
    if rule is None:
        rule = np.where(syn.plastic != 0, 0, -1) if syn.plastic is not None else -1
    recs = records(pre, syn.post, syn.type, syn.weight, syn.delay, rule=rule)
    perm = np.random.default_rng(seed).permutation(len(recs))
    write_nbnet(path, META, recs[perm])


def test_loaded_network_matches_oracle(gpu, tmp_path):
    wl = small_sparse(L.LIF_VJ, L.RK4)
    path = str(tmp_path / "lif.nbnet")
    file_from_csr(path, wl.syn)
    net = wl.make_net(nbnet=path)
    # rows sorted by (delay, post): the device sort restores the workload's CSR exactly
    assert_bits_equal(net.weights(), wl.syn.weight, "weights in CSR order")
    T = 300
    net.record(0, T, counts=True, raster=True)
    net.run(T, L.ENGINE_PERSISTENT)
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    counts_o, raster_o = run.run(T)
    assert counts_o.sum() > 100
    compare(net, run, wl, counts_o, raster_o)


def test_loaded_plastic_network_matches_oracle(gpu, tmp_path):
    # synapse_rule >= 0 marks the plastic synapses; STDP weights and traces stay bit-exact
    wl = W.stdp_izh(n=1500, k_in=60, steps=600, w_bg=2.0)
    path = str(tmp_path / "stdp.nbnet")
    file_from_csr(path, wl.syn, seed=3)
    net = wl.make_net(nbnet=path)
    T = 600
    net.record(0, T, counts=True, raster=True)
    net.run(T, L.ENGINE_GRAPH)
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    counts_o, raster_o = run.run(T)
    assert counts_o.sum() > 100
    compare(net, run, wl, counts_o, raster_o, ring=False)
    w = net.weights()
    assert_bits_equal(w, run.w, "STDP weights")
    apre, apost, last = net.stdp_state()
    assert_bits_equal(apre, run.apre, "Apre")
    assert_bits_equal(apost, run.apost, "Apost")
    assert_bits_equal(last, run.slast, "LastUpdate")
    assert not np.array_equal(w, wl.syn.weight), "weights never changed"


@pytest.mark.parametrize("field,value,status", [
    ("delay", 0, L.NBX_ERR_UNSUPPORTED),
    ("delay", 64, L.NBX_ERR_UNSUPPORTED),
    ("type", 2, L.NBX_ERR_UNSUPPORTED),
    ("rule", -2, L.NBX_ERR_INVALID),
    ("trigger", -5, L.NBX_ERR_INVALID),
    ("dst", 10_000, L.NBX_ERR_INVALID),
    ("weight", np.inf, L.NBX_ERR_INVALID),
])
def test_load_rejects(gpu, tmp_path, field, value, status):
    wl = small_sparse(L.LIF_VJ, L.EULER, n=500, k_in=20)
    n = wl.spec.n
    pre = np.repeat(np.arange(n), np.diff(wl.syn.row_ptr))
    recs = records(pre, wl.syn.post, wl.syn.type, wl.syn.weight, wl.syn.delay)
    recs[17][field] = value
    path = str(tmp_path / "bad.nbnet")
    write_nbnet(path, META, recs)
    from neuronbridge_amd.network import Net
    net = Net(wl.spec)
    with pytest.raises(L.NbxError) as e:
        net.load_nbnet(path)
    assert e.value.status == status


def test_load_after_finalize_is_rejected(gpu, tmp_path):
    wl = small_sparse(L.LIF_VJ, L.EULER, n=500, k_in=20)
    path = str(tmp_path / "ok.nbnet")
    file_from_csr(path, wl.syn)
    net = wl.make_net()
    with pytest.raises(L.NbxError) as e:
        net.load_nbnet(path)
    assert e.value.status == L.NBX_ERR_STATE


def test_large_table_in_several_chunks(gpu, tmp_path):
    # 1e7 records (three device transfers): the loaded network equals the in-memory one
    wl = W.sparse_lif(n=100_000, k_in=100)
    path = str(tmp_path / "big.nbnet")
    file_from_csr(path, wl.syn, seed=5)
    a = wl.make_net(nbnet=path)
    b = wl.make_net()
    assert_bits_equal(a.weights(), b.weights(), "weights")
    T = 200
    for net in (a, b):
        net.record(0, T, counts=True, raster=True)
        net.run(T, L.ENGINE_PERSISTENT)
    assert_bits_equal(a.raster(), b.raster(), "raster")
    assert_bits_equal(a.ring(), b.ring(), "ring")
    assert a.events() == b.events()


def test_network_from_nbnet_metadata_matches_oracle(gpu, tmp_path):
    # a network described by a .nbnet file alone: two layers (LIF voltage jump, conductance
    # Izhikevich with NMDA synapses) with the reference's parameter names, tagged and plain values and
    # initial-state offsets; populations, state and output mask come from the metadata
    from neuronbridge_amd import nbnet_meta as M
    from neuronbridge_amd.workloads import Workload
    base = W.populations((L.LIF_VJ, L.IZH), (1600, 400), k_in=100, integrator=L.EULER)
    tag = lambda kind, v: {M.MARKER: kind, "value": v}
    meta = dict(META)
    meta["layers"] = [
        {"model": "TimeDrivenLIF_Voltage_jump", "count": 1600, "output": True, "parameters": {
            "V_rest": -65.0, "tau": tag("float32", 20.0), "V_th": -50.0, "R": 1.0, "V_reset": -65.0,
            "t_ref": tag("int", 20), "random_mu": tag("float32_list", [8.0, 0.0])}},
        {"model": "TimeDrivenIzhikevic_Exponential_Decay", "count": 400, "monitored": True, "parameters": {
            "a": 0.1, "b": 0.2, "c": -65.0, "d": 2.0, "V_rest": -65.0, "nmda_tau": 100.0,
            "random_mu": [5.0, 0.0, 0.0, 0.0, 0.0, 0.0], "random_sigma": [0, 0, 0, 0, 0, 0]}}]
    n = 2000
    syn = base.syn
    pre = np.repeat(np.arange(n, dtype=np.int64), np.diff(syn.row_ptr))
    recs = records(pre, syn.post, syn.type, syn.weight, syn.delay)
    path = str(tmp_path / "meta.nbnet")
    write_nbnet(path, meta, recs[np.random.default_rng(3).permutation(len(recs))])
    net, desc = M.load_network(path, dt=0.1, integrator=L.EULER, poisson=(base.pois_seed, 1.0, 0.1))
    assert [ly.count for ly in desc.layers] == [1600, 400] and list(desc.monitored) == list(range(1600, 2000))
    wl = Workload("from_metadata", desc.spec, syn=syn, pois_seed=base.pois_seed, pois_lambda=1.0, w_bg=0.1,
                  x0=desc.x0, last0=desc.last0)
    x, _ = net.get_state()
    assert_bits_equal(x, desc.x0, "initial state from the metadata")
    T = 300
    net.record(0, T, counts=True, raster=True)
    net.run(T, L.ENGINE_GRAPH)
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    counts_o, raster_o = run.run(T)
    bits = np.unpackbits(raster_o.view(np.uint8), axis=1, bitorder="little")[:, :n]
    assert bits[:, :1600].sum() > 0 and bits[:, 1600:].sum() > 0, "a layer is silent"
    compare(net, run, wl, counts_o, raster_o)
    # the output mask of the layer marked output: one more per-phase step
    net.begin_step(T)
    net.run_propagation()
    out, full = net.neuron_step_outputs()
    assert set(out.tolist()) == {i for i in full.tolist() if i < 1600}
    net.commit_firings()
    net.close()
