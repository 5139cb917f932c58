"""GPU parity of the trigger-driven learning rules (R-STDP eligibility with reward / punishment
triggers, additive kernel) against the oracle: weights, traces, eligibilities and spikes bit-exact."""
import numpy as np
import pytest

from harness import OracleRun
from neuronbridge_amd import _lib as L
from neuronbridge_amd import workloads as W
from test_gpu_parity import ENGINES, assert_bits_equal, compare


def check_rules(net, run, wl):
    sel = wl.rules["rule"] >= 0
    assert_bits_equal(net.weights(), run.w, "weights")
    apre, apost, elig, last = net.rule_state()
    assert_bits_equal(apre[sel], run.apre[sel], "Apre")
    assert_bits_equal(apost[sel], run.apost[sel], "Apost")
    assert_bits_equal(elig[sel], run.elig[sel], "eligibility")
    assert_bits_equal(last[sel], run.slast[sel], "last update")


pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("with_stdp,cerebellar", [(False, False), (True, False), (False, True)],
                         ids=["rules", "rules+stdp", "rules+cerebellar"])
@pytest.mark.parametrize("engine", ["kernels", "graph"])
def test_rule_network_bitexact(gpu, engine, with_stdp, cerebellar):
    wl = W.trigger_rules(with_stdp=with_stdp, cerebellar=cerebellar)
    T = 400
    net = wl.make_net()
    w0 = net.weights()
    net.record(0, T, counts=True, raster=True)
    net.run(T, ENGINES[engine])
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    counts_o, raster_o = run.run(T)
    assert counts_o.sum() > 500
    compare(net, run, wl, counts_o, raster_o, ring=False)
    check_rules(net, run, wl)
    rule = wl.rules["rule"]
    assert np.mean(net.weights()[rule == 0] != w0[rule == 0]) > 0.05, "R-STDP never triggered"
    assert np.mean(net.weights()[rule == 1] != w0[rule == 1]) > 0.5
    if cerebellar:
        assert np.mean(net.weights()[rule == 2] != w0[rule == 2]) > 0.5
    if with_stdp:
        assert_bits_equal(net.stdp_state()[0][wl.syn.plastic != 0], run.apre[wl.syn.plastic != 0], "STDP Apre")


def test_many_simultaneous_triggers(gpu):
    # 30% of the synapses are triggers: targets often take several trigger arrivals in one step,
    # which must apply in ascending synapse order
    wl = W.trigger_rules(n=1500, k_in=80, seed=3, trigger_fraction=0.3)
    T = 300
    net = wl.make_net()
    net.record(0, T, counts=True, raster=True)
    net.run(T, L.ENGINE_GRAPH)
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    counts_o, raster_o = run.run(T)
    compare(net, run, wl, counts_o, raster_o, ring=False)
    check_rules(net, run, wl)


def test_rules_through_per_phase_calls(gpu):
    # the reference's per-step call sequence drives the same trigger and post-side phases
    wl = W.trigger_rules(n=1200, k_in=60, seed=5)
    T = 120
    net = wl.make_net()
    net.record(0, T, counts=True, raster=True)
    for t in range(T):
        net.begin_step(t)
        net.run_propagation()
        net.neuron_step()
        net.apply_post_learning()
        net.commit_firings()
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    counts_o, raster_o = run.run(T)
    compare(net, run, wl, counts_o, raster_o, ring=False)
    check_rules(net, run, wl)


def test_rule_network_refuses_persistent_engine(gpu):
    wl = W.trigger_rules(n=800, k_in=40)
    net = wl.make_net()
    with pytest.raises(L.NbxError) as e:
        net.run(3, L.ENGINE_PERSISTENT)
    assert e.value.status == L.NBX_ERR_UNSUPPORTED
    assert net.replay_path() == L.REPLAY_GLOBAL


def test_nbnet_learning_rules(gpu, tmp_path):
    # a .nbnet table whose records name rules 0 (R-STDP), 1 (additive) and 2 (pair STDP) and carry
    # per-record weight bounds; trigger ids naming the STDP rule are ignored like the reference does
    from nbnet_file import records, write_nbnet
    wl = W.trigger_rules(n=1500, k_in=80, seed=9, with_stdp=True)
    s = wl.syn
    S = s.count
    rng = np.random.default_rng(4)
    maxw = rng.uniform(0.2, 0.4, S).astype(np.float32)
    wl.rules["max_weight"] = maxw
    wl.stdp["max_weight"] = maxw
    rule_id = np.where(s.plastic != 0, 2, wl.rules["rule"]).astype(np.int32)
    trig_id = wl.rules["trigger"].copy()
    static = (rule_id < 0) & (trig_id < 0)
    trig_id[static & (rng.random(S) < 0.05)] = 2  # triggers of a rule without a trigger side
    n = wl.spec.n
    pre = np.repeat(np.arange(n, dtype=np.int64), np.diff(s.row_ptr))
    recs = records(pre, s.post, s.type, s.weight, s.delay, max_weight=maxw, rule=rule_id, trigger=trig_id)
    recs = recs[rng.permutation(S)]
    path = str(tmp_path / "rules.nbnet")
    write_nbnet(path, {"learning_rules": ["R_STDP", "AdditiveKernalChange", "STDP"]}, recs)
    st = wl.stdp
    stdp_rule = dict(kind=L.RULE_STDP, max_ltp=st["max_ltp"], ltp_tau=st["ltp_tau"], max_ltd=st["max_ltd"],
                     ltd_tau=st["ltd_tau"])
    net = wl.make_net(nbnet=path, nbnet_rules=[wl.rules["rules"][0], wl.rules["rules"][1], stdp_rule])
    T = 300
    net.record(0, T, counts=True, raster=True)
    net.run(T, L.ENGINE_GRAPH)
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    counts_o, raster_o = run.run(T)
    compare(net, run, wl, counts_o, raster_o, ring=False)
    check_rules(net, run, wl)
    plastic = s.plastic != 0
    assert np.any(run.w[plastic] != s.weight[plastic])


def test_nbnet_rule_ids_are_checked(gpu, tmp_path):
    from nbnet_file import records, write_nbnet
    wl = W.trigger_rules(n=600, k_in=20, seed=2)
    s = wl.syn
    n = wl.spec.n
    pre = np.repeat(np.arange(n, dtype=np.int64), np.diff(s.row_ptr))
    path = str(tmp_path / "r.nbnet")
    write_nbnet(path, {}, records(pre, s.post, s.type, s.weight, s.delay, rule=wl.rules["rule"],
                                  trigger=wl.rules["trigger"]))
    with pytest.raises(L.NbxError) as e:  # the file names rules 0 and 1; only one is given
        wl.make_net(nbnet=path, nbnet_rules=[wl.rules["rules"][0]])
    assert e.value.status == L.NBX_ERR_INVALID
    bad = records(pre, s.post, s.type, s.weight, s.delay, rule=-2)
    write_nbnet(path, {}, bad)
    with pytest.raises(L.NbxError) as e:
        wl.make_net(nbnet=path)
    assert e.value.status == L.NBX_ERR_INVALID


def test_cerebellar_history_overflow_is_reported(gpu):
    # one arrival kept per synapse while the kernel window spans many arrivals: the run stops with
    # an error instead of silently dropping arrivals the next trigger would need
    wl = W.trigger_rules(n=1200, k_in=60, seed=6, cerebellar=True, history=1)
    net = wl.make_net()
    with pytest.raises(L.NbxError) as e:
        net.run(600, L.ENGINE_GRAPH)
    assert e.value.status == L.NBX_ERR_STATE and "history" in str(e.value)


def test_cerebellar_rule_from_nbnet_uses_record_order(gpu, tmp_path):
    # the cerebellar random stream is keyed by the connection's index in the file
    from nbnet_file import records, write_nbnet
    wl = W.trigger_rules(n=1200, k_in=60, seed=8, cerebellar=True)
    s = wl.syn
    n = wl.spec.n
    pre = np.repeat(np.arange(n, dtype=np.int64), np.diff(s.row_ptr))
    recs = records(pre, s.post, s.type, s.weight, s.delay, max_weight=wl.rules["max_weight"],
                   rule=wl.rules["rule"], trigger=wl.rules["trigger"])
    perm = np.random.default_rng(3).permutation(s.count)
    path = str(tmp_path / "cb.nbnet")
    write_nbnet(path, {}, recs[perm])
    conn = np.empty(s.count, np.int32)
    conn[perm] = np.arange(s.count, dtype=np.int32)  # CSR synapse j is record conn[j]
    wl.extra["conn_index"] = conn
    net = wl.make_net(nbnet=path, nbnet_rules=wl.rules["rules"])
    T = 300
    net.record(0, T, counts=True, raster=True)
    net.run(T, L.ENGINE_GRAPH)
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    counts_o, raster_o = run.run(T)
    compare(net, run, wl, counts_o, raster_o, ring=False)
    check_rules(net, run, wl)
