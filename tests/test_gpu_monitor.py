"""Monitor export in the reference's DebugMonitor directory schema
(src/native/gpu/src/debug_monitor/DebugMonitorWriter.cpp:193-249,
MainNetworkDebugSource.cpp:97-145, python/src/neuronbridge/results.py readers)."""
import csv
import json
import os

import numpy as np
import pytest

from nbnet_file import records, write_nbnet
from neuronbridge_amd import _lib as L
from test_gpu_parity import small_sparse

pytestmark = pytest.mark.gpu

HEADERS = {
    "neuron_state.csv": "time_step,component_kind,component_index,component_name,global_neuron_id,"
                        "local_neuron_id,field_name,value",
    "spikes.csv": "time_step,component_kind,component_index,component_name,global_neuron_id,local_neuron_id",
    "pending_channels.csv": "time_step,subnetwork_index,subnetwork_name,channel,local_neuron_id,global_neuron_id,value",
    "inputconv_outputs.csv": "time_step,inputconv_index,inputconv_name,output_index,value",
    "inputconv_inputs.csv": "time_step,inputconv_index,inputconv_name,input_index,value",
    "inputconv_state.csv": "time_step,inputconv_index,inputconv_name,field_name,state_index,value",
    "outer_dynamic_state.csv": "time_step,component_index,component_name,field_name,index,value",
    "weights.csv": "time_step,component_kind,component_index,component_name,synapse_id,value",
}
META_KEYS = ["sample_interval_steps", "flush_interval_steps", "record_state", "record_spikes", "record_weights",
             "record_pending_channels", "record_inputconv_outputs", "record_outer_dynamic_state"]


def rows(d, name):
    with open(os.path.join(d, name), newline="") as f:
        return list(csv.DictReader(f))


def close(got, want):
    # values are written with the stream's default six significant digits, as the reference writes them
    np.testing.assert_allclose(np.asarray(got, np.float64), np.asarray(want, np.float64), rtol=5e-6, atol=1e-30)


@pytest.mark.parametrize("model", [L.LIF_VJ, L.IZH])
def test_export_schema_and_values(gpu, tmp_path, model):
    wl = small_sparse(model, L.RK4 if model == L.IZH else L.EULER)
    net = wl.make_net()
    T, every = 120, 3
    net.record(0, T, counts=False, raster=True)
    net.run(T)
    sel = np.array([0, 5, 17, 400, wl.spec.n - 1], dtype=np.int32)
    out = str(tmp_path / "monitor")
    net.export_monitor(out, sample_interval=every, flush_interval=50, weights=True, neuron_ids=sel, name="main")

    for name, header in HEADERS.items():
        with open(os.path.join(out, name)) as f:
            assert f.readline().rstrip("\n") == header, name
    with open(os.path.join(out, "meta.json")) as f:
        meta = json.load(f)
    assert list(meta) == META_KEYS
    assert meta["sample_interval_steps"] == every and meta["flush_interval_steps"] == 50
    assert meta["record_state"] and meta["record_spikes"] and meta["record_weights"]
    assert not meta["record_inputconv_outputs"] and not meta["record_outer_dynamic_state"]

    # spikes: every recorded step divisible by the interval, selected neurons only
    r = net.raster()
    want = sorted((t, int(i)) for t in range(0, T, every) for i in sel if (r[t, i >> 5] >> (i & 31)) & 1)
    sp = rows(out, "spikes.csv")
    got = sorted((int(x["time_step"]), int(x["global_neuron_id"])) for x in sp)
    assert got == want and len(want) > 0
    assert all(x["component_kind"] == "MainNetwork" and x["component_index"] == "0" and
               x["component_name"] == "main" and x["local_neuron_id"] == x["global_neuron_id"] for x in sp)

    # state: nstate variables, then the step of the last update and two entries reading -1
    x, _ = net.get_state()
    ns = x.shape[1]
    st = rows(out, "neuron_state.csv")
    assert len(st) == sel.size * (ns + 3)
    fields = ["v"] + [f"state_{k}" for k in range(1, ns + 3)]
    for k, i in enumerate(sel):
        mine = st[k * (ns + 3):(k + 1) * (ns + 3)]
        assert [m["field_name"] for m in mine] == fields
        assert all(int(m["time_step"]) == T - 1 and int(m["global_neuron_id"]) == i for m in mine)
        vals = [float(m["value"]) for m in mine]
        close(vals[:ns], x[i])
        assert vals[ns:] == [T - 1, -1.0, -1.0]

    # pending arrivals of the next step, from the delay ring
    ring = net.ring()
    scale = 2.0 ** -net.fixed_shift
    pend = rows(out, "pending_channels.csv")
    nch = ring.shape[1]
    assert len(pend) == nch * sel.size
    for p in pend:
        assert int(p["time_step"]) == T
        close(float(p["value"]), ring[T % ring.shape[0], int(p["channel"]), int(p["local_neuron_id"])] * scale)

    # weights in CSR order for a network built from arrays
    w = net.weights()
    wr = rows(out, "weights.csv")
    assert [int(x["synapse_id"]) for x in wr] == list(range(w.size))
    close([float(x["value"]) for x in wr], w)


def test_export_weights_follow_file_order(gpu, tmp_path):
    wl = small_sparse(L.LIF_VJ, L.EULER, n=600, k_in=25)
    n = wl.spec.n
    pre = np.repeat(np.arange(n, dtype=np.int64), np.diff(wl.syn.row_ptr))
    recs = records(pre, wl.syn.post, wl.syn.type, wl.syn.weight, wl.syn.delay)
    recs = recs[np.random.default_rng(7).permutation(len(recs))]
    path = str(tmp_path / "net.nbnet")
    write_nbnet(path, {"layers": []}, recs)
    net = wl.make_net(nbnet=path)
    net.run(10)
    out = str(tmp_path / "mon")
    net.export_monitor(out, spikes=False, state=False, pending=False, weights=True)
    wr = rows(out, "weights.csv")
    assert [int(x["synapse_id"]) for x in wr] == list(range(len(recs)))
    close([float(x["value"]) for x in wr], recs["weight"])
    assert rows(out, "spikes.csv") == [] and rows(out, "neuron_state.csv") == []


def test_export_rejects(gpu, tmp_path):
    wl = small_sparse(L.LIF_VJ, L.EULER, n=300, k_in=10)
    net = wl.make_net()
    with pytest.raises(L.NbxError) as e:
        net.export_monitor(str(tmp_path / "a"), sample_interval=0)
    assert e.value.status == L.NBX_ERR_INVALID
    with pytest.raises(L.NbxError) as e:
        net.export_monitor(str(tmp_path / "b"), neuron_ids=[300])
    assert e.value.status == L.NBX_ERR_INVALID
    net.begin_step(0)
    with pytest.raises(L.NbxError) as e:
        net.export_monitor(str(tmp_path / "c"))
    assert e.value.status == L.NBX_ERR_STATE


# ---------------------------------------------------------------- time series (SimulationDebugMonitor)

def _oracle_frames(wl, shift, steps, last_step):
    """The oracle stepped to each sampled step: {s: (state rows, raster row, ring)}."""
    from harness import OracleRun
    run = OracleRun(wl, fixed_shift=shift)
    frames = {}
    for s in steps:
        _, ras = run.run(s + 1 - run.t)
        frames[s] = (run.x.copy(), ras[-1].copy(), run.ring_as_device(wl.spec.channels).copy())
    if run.t < last_step:
        run.run(last_step - run.t)
    return frames


@pytest.mark.parametrize("engine", ["graph", "persistent", "per_phase"])
def test_time_series_monitor_against_oracle(gpu, tmp_path, engine):
    wl = small_sparse(L.LIF_VJ, L.RK4)
    net = wl.make_net()
    n, every, T = wl.spec.n, 7, 100
    sel = np.arange(1, n, 4, dtype=np.int32)
    out = str(tmp_path / "ts")
    net.monitor_start(out, sample_interval=every, flush_interval=4, neuron_ids=sel, name="main")
    if engine == "per_phase":
        for t in range(T):
            net.begin_step(t)
            net.run_propagation()
            net.neuron_step()
            net.commit_firings()
    else:
        eng = L.ENGINE_GRAPH if engine == "graph" else L.ENGINE_PERSISTENT
        net.run(45, eng)   # the sampled steps fall inside and at the end of runs
        net.run(T - 45, eng)
    written, held = net.monitor_frames()
    steps = list(range(0, T, every))
    assert written + held == len(steps) and written == 4 * (len(steps) // 4)
    net.monitor_stop()
    assert net.monitor_frames() == (0, 0)

    with open(os.path.join(out, "meta.json")) as f:
        meta = json.load(f)
    assert list(meta) == META_KEYS and meta["sample_interval_steps"] == every and meta["flush_interval_steps"] == 4
    for name, header in HEADERS.items():
        with open(os.path.join(out, name)) as f:
            assert f.readline().rstrip("\n") == header, name

    frames = _oracle_frames(wl, net.fixed_shift, steps, T)
    ns = wl.spec.nstate
    # spikes: the oracle's raster at every sampled step, selected neurons, ascending within a step
    want = [(s, int(i)) for s in steps for i in sel if (frames[s][1][i >> 5] >> (i & 31)) & 1]
    got = [(int(r["time_step"]), int(r["global_neuron_id"])) for r in rows(out, "spikes.csv")]
    assert got == want and len(want) > 0
    # state after each sampled step's update
    st = rows(out, "neuron_state.csv")
    assert len(st) == len(steps) * sel.size * (ns + 3)
    k = 0
    for s in steps:
        x = frames[s][0]
        for i in sel:
            mine = st[k:k + ns + 3]
            k += ns + 3
            assert all(int(m["time_step"]) == s and int(m["global_neuron_id"]) == i for m in mine)
            vals = [float(m["value"]) for m in mine]
            close(vals[:ns], x[i])
            assert vals[ns:] == [s, -1.0, -1.0]
    # arrivals pending for step s + 1
    pend = rows(out, "pending_channels.csv")
    nch = frames[steps[0]][2].shape[1]
    assert len(pend) == len(steps) * nch * sel.size
    scale = 2.0 ** -net.fixed_shift
    for p in pend:
        s, ch, i = int(p["time_step"]), int(p["channel"]), int(p["local_neuron_id"])
        ring = frames[s][2]
        close(float(p["value"]), ring[(s + 1) % ring.shape[0], ch, i] * scale)
    assert rows(out, "weights.csv") == []
    # the run itself is unchanged by the monitor
    x, _ = net.get_state()
    assert np.array_equal(x.view(np.uint32), _final_state(wl, net.fixed_shift, T).view(np.uint32))
    net.close()


def _final_state(wl, shift, T):
    from harness import OracleRun
    run = OracleRun(wl, fixed_shift=shift)
    run.run(T)
    return run.x


def test_time_series_monitor_stdp_weights_change(gpu, tmp_path):
    # a plastic network: the weight rows of later frames differ from the first and equal the
    # oracle's weights after the same steps
    from harness import OracleRun
    from neuronbridge_amd import workloads as W
    wl = W.stdp_izh(n=1500, k_in=60, steps=400, w_bg=2.0)
    net = wl.make_net()
    out = str(tmp_path / "stdp")
    net.monitor_start(out, sample_interval=100, flush_interval=100, spikes=False, state=False, pending=False,
                      weights=True)
    net.run(400, L.ENGINE_GRAPH)
    net.monitor_stop()
    wr = rows(out, "weights.csv")
    S = wl.syn.count
    assert len(wr) == 4 * S
    assert [int(r["time_step"]) for r in wr[::S]] == [0, 100, 200, 300]
    assert [int(r["synapse_id"]) for r in wr[:S]] == list(range(S))
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    for f, s in enumerate((0, 100, 200, 300)):
        run.run(s + 1 - run.t)
        vals = np.array([float(r["value"]) for r in wr[f * S:(f + 1) * S]])
        close(vals, run.w)
    assert (np.array([float(r["value"]) for r in wr[3 * S:]]) != wl.syn.weight).any()


def test_time_series_monitor_rejects(gpu, tmp_path):
    wl = small_sparse(L.LIF_VJ, L.EULER, n=300, k_in=10)
    net = wl.make_net()
    with pytest.raises(L.NbxError) as e:
        net.monitor_start(str(tmp_path / "a"), sample_interval=0)
    assert e.value.status == L.NBX_ERR_INVALID
    net.monitor_start(str(tmp_path / "b"))
    with pytest.raises(L.NbxError) as e:
        net.monitor_start(str(tmp_path / "c"))
    assert e.value.status == L.NBX_ERR_STATE
    net.run(3)
    net.monitor_stop()
    net.monitor_stop()  # stopping twice is harmless
    with pytest.raises(L.NbxError) as e:
        net.monitor_flush()
    assert e.value.status == L.NBX_ERR_STATE
    assert len(rows(str(tmp_path / "b"), "spikes.csv")) >= 0
    assert len(rows(str(tmp_path / "b"), "neuron_state.csv")) == 3 * 300 * (wl.spec.nstate + 3)
