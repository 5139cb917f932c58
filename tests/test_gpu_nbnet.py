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
