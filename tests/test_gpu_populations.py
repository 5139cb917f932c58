"""Networks of several populations, the Izhikevich NMDA channel, external inputs (interface
channels, arrivals from outside the synapses), output masks and ResetState on the GPU, bit for bit
against the CPU oracle.

The interface records follow the reference's semantics (GpuPropagationRuntime.cu:68-140,
DenseSubnetworkRuntimeGpu.cpp:625-663): a record names its target directly or through an interface
slot, zero host values are skipped, value * scale is added to the step's pending input of the
record's channel (excitatory / inhibitory / NMDA as synapse types 0 / 1 / 2), current records
replace I_ext. The expected inputs handed to the oracle are resolved here from the same records.
"""
import copy

import numpy as np
import pytest
import torch  # noqa: F401  (torch's HIP runtime first)

import oracle as O
from harness import OracleRun
from neuronbridge_amd import _lib as L
from neuronbridge_amd import workloads as W
from neuronbridge_amd.hiptime import DeviceBuffer
from neuronbridge_amd.network import ModelSpec
from test_gpu_parity import ENGINES, assert_bits_equal, compare, small_sparse

pytestmark = pytest.mark.gpu

MIXES = {
    "lif_izh": ((L.LIF_VJ, L.IZH), (1600, 400)),
    "izhcur_exp3_poisson_lifexp": ((L.IZH_CUR, L.LIF_EXP3, L.POISSON, L.LIF_EXP), (700, 500, 300, 500)),
    "exp2_relay_lif": ((L.LIF_EXP2, L.TRIGGER_RELAY, L.LIF_VJ), (900, 100, 1000)),
}


@pytest.mark.parametrize("mix", list(MIXES))
@pytest.mark.parametrize("engine", ["kernels", "graph"])
@pytest.mark.parametrize("integrator", [L.EULER, L.RK4], ids=["euler", "rk4"])
def test_populations_closed_loop_bitexact(gpu, mix, engine, integrator):
    models, sizes = MIXES[mix]
    wl = W.populations(models, sizes, k_in=100, integrator=integrator)
    T = 300
    net = wl.make_net()
    assert net.pick_engine() == L.ENGINE_GRAPH
    net.record(0, T, counts=True, raster=True)
    net.run(T, ENGINES[engine])
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    counts_o, raster_o = run.run(T)
    bits = np.unpackbits(raster_o.view(np.uint8), axis=1, bitorder="little")[:, :wl.spec.n]
    b = wl.extra["begins"]
    for k in range(len(sizes)):
        assert bits[:, b[k]:b[k + 1]].sum() > 0, f"population {k} silent"
    compare(net, run, wl, counts_o, raster_o)


def test_populations_persistent_engine_rejected(gpu):
    wl = W.populations(k_in=50)
    net = wl.make_net()
    with pytest.raises(L.NbxError, match="several populations"):
        net.run(5, L.ENGINE_PERSISTENT)
    net.run(5, L.ENGINE_AUTO)


def izh_nmda_workload(integrator, n=2000, k_in=100, seed=44):
    wl = small_sparse(L.IZH, integrator, n=n, k_in=k_in, seed=seed)
    rng = np.random.default_rng(seed)
    wl.spec.nmda_tau = 100.0
    nmda = (wl.syn.type == 0) & (rng.random(wl.syn.count) < 0.3)
    wl.syn.type = np.where(nmda, 2, wl.syn.type).astype(np.uint8)
    wl.syn.weight = np.where(nmda, np.float32(0.04), wl.syn.weight).astype(np.float32)
    return wl


@pytest.mark.parametrize("engine", ["kernels", "graph"])
@pytest.mark.parametrize("integrator", [L.EULER, L.RK4], ids=["euler", "rk4"])
def test_izhikevich_nmda_bitexact(gpu, engine, integrator):
    # TimeDrivenIzhikevic_Exponential_Decay.cpp:69-72,96-97,127-128: type-2 synapses into g_nmda
    wl = izh_nmda_workload(integrator)
    T = 300
    net = wl.make_net()
    assert net.pick_engine() == L.ENGINE_GRAPH
    net.record(0, T, counts=True, raster=True)
    net.run(T, ENGINES[engine])
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    counts_o, raster_o = run.run(T)
    assert counts_o.sum() > 50
    assert np.abs(run.x[:, 4]).max() > 0, "no NMDA conductance built up"
    compare(net, run, wl, counts_o, raster_o, ring=False)
    assert_bits_equal(net.ring(), run.ring_as_device(3), "delay ring (AMPA, GABA, NMDA channels)")


def _resolve(records, slots, n_models, model_of, kind_of):
    """Reference semantics of one step's host records -> oracle inputs (target, kind, value)."""
    out = []
    for slot, tgt, ch, val, sc, ow in records:
        if np.float32(val) == 0.0:
            continue
        t = tgt if tgt >= 0 else (slots[slot] if 0 <= slot < len(slots) else -1)
        if t < 0:
            continue
        kind = kind_of(ch, model_of(t))
        if kind is None:
            continue
        out.append((t, kind, np.float32(val) * np.float32(sc)))
    return out


def _kind_of(ch, model):
    if ch == L.PENDING_CURRENT:
        return None if model == L.TRIGGER_RELAY else O.INPUT_CURRENT
    kind = {L.PENDING_EXC: 0, L.PENDING_INH: 1, L.PENDING_NMDA: 2}.get(ch)
    if kind is None:
        return None
    max_type = 2 if model in (L.IZH, L.LIF_EXP3) else 1
    return kind if kind <= max_type else None


@pytest.mark.parametrize("which", ["one_population", "populations"])
def test_interface_inputs_and_arrivals_bitexact(gpu, which):
    if which == "populations":
        wl = W.populations((L.LIF_VJ, L.IZH), (1200, 400), k_in=80)
        begins = wl.extra["begins"]
        models = [p.model for p in wl.spec.pops]
        model_of = lambda i: models[int(np.searchsorted(begins, i, side="right") - 1)]  # noqa: E731
    else:
        wl = small_sparse(L.LIF_VJ, L.EULER, n=1600, k_in=80, seed=8)
        model_of = lambda i: L.LIF_VJ  # noqa: E731
    n = wl.spec.n
    T = 120
    rng = np.random.default_rng(17)
    nslots = 64
    slots = rng.integers(0, n, nslots).astype(np.int32)
    slots[::9] = -1  # unmapped slots are skipped
    slot_ch = rng.choice([L.PENDING_EXC, L.PENDING_INH, L.PENDING_NMDA], nslots)
    cur_slots = [3, 20, 41]  # current slots (overwrite)
    for s in cur_slots:
        slot_ch[s] = L.PENDING_CURRENT
    nsrc = 24
    copy_specs = [(-1, int(rng.integers(0, n)), int(rng.choice([L.PENDING_EXC, L.PENDING_NMDA, L.PENDING_INH])),
                   int(rng.integers(0, nsrc)), 0.25, False) for _ in range(30)]
    copy_specs.append((5, -1, L.PENDING_EXC, 2, 2.0, True))  # through a slot, overwrite
    d_vals = DeviceBuffer(4 * nsrc)
    net = wl.make_net()
    net.set_interface_targets(slots)
    net.record(0, T, counts=True, raster=True)
    inputs = [[] for _ in range(T + 8)]
    for t in range(T):
        vals = (rng.random(nslots) < 0.4) * rng.uniform(0.0, 3.0, nslots)
        recs = []
        for s in range(nslots):
            if slot_ch[s] == L.PENDING_CURRENT:
                recs.append((s, -1, L.PENDING_CURRENT, float(rng.uniform(0.0, 6.0)) if t % 7 else 0.0, 1.0, True))
            else:
                recs.append((s, -1, int(slot_ch[s]), float(vals[s]), 0.5, False))
        src = np.where(rng.random(nsrc) < 0.5, rng.uniform(0.0, 2.0, nsrc), 0.0).astype(np.float32)
        d_vals.upload(src)
        net.begin_step(t)
        net.upload_interface_values(recs)
        net.apply_interface()
        net.upload_interface_values_device(d_vals.ptr, copy_specs)
        # host records: zero values skipped except that current records set I_ext (overwrite)
        step_in = []
        for r in recs:
            if r[2] == L.PENDING_CURRENT and r[3] == 0.0:
                continue  # the reference's kernel skips a zero value, current records included
            step_in += _resolve([r], slots, None, model_of, _kind_of)
        for slot, tgt, ch, si, sc, ow in copy_specs:
            v = np.float32(src[si]) * np.float32(sc)
            if v == 0.0 and not ow:
                continue
            tt = tgt if tgt >= 0 else int(slots[slot])
            if tt < 0:
                continue
            kind = _kind_of(ch, model_of(tt))
            if kind is not None:
                step_in.append((tt, kind, v))
        inputs[t] += step_in
        # arrivals from outside the synapses, 0..3 steps ahead (none still pending at the end)
        if t < T - 4:
            k = int(rng.integers(1, 6))
            tg = rng.integers(0, n, k)
            ty = np.array([min(int(rng.integers(0, 3)), 2 if model_of(int(i)) in (L.IZH,) else 1) for i in tg],
                          np.uint8)
            w = rng.uniform(0.0, 2.0, k).astype(np.float32)
            ahead = int(rng.integers(0, 4))
            net.add_arrivals(tg, ty, w, t + ahead)
            inputs[t + ahead] += [(int(a), int(b_), float(c)) for a, b_, c in zip(tg, ty, w)]
        net.run_propagation()
        net.neuron_step()
        net.commit_firings()
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    run.set_inputs(inputs, 0)
    counts_o, raster_o = run.run(T)
    assert counts_o.sum() > 50
    compare(net, run, wl, counts_o, raster_o)


def test_interface_rejections(gpu):
    wl = small_sparse(L.LIF_VJ, L.EULER, n=500, k_in=50, seed=3)
    net = wl.make_net()
    net.set_interface_targets(np.arange(10, dtype=np.int32))
    with pytest.raises(L.NbxError, match="overwrite"):
        net.upload_interface_values([(0, -1, L.PENDING_CURRENT, 1.0, 1.0, False)] * 10)
    with pytest.raises(L.NbxError, match="mapping count"):
        net.upload_interface_values([(0, -1, L.PENDING_EXC, 1.0, 1.0, False)] * 3)
    with pytest.raises(L.NbxError, match="overflow"):
        net.upload_interface_values([(0, -1, L.PENDING_EXC, 1e9, 1.0, False)] * 10)
        net.apply_interface()
    with pytest.raises(L.NbxError, match="reach"):
        net.add_arrivals([1], [0], [0.5], net.time + 100)
    net.begin_step(net.time)
    net.run_propagation()
    net.neuron_step()
    with pytest.raises(L.NbxError, match="before its neuron update"):
        net.apply_interface()
    net.commit_firings()


def test_output_mask_and_pending_download(gpu):
    wl = W.populations((L.LIF_VJ, L.IZH), (800, 200), k_in=60)
    net = wl.make_net()
    mask = np.zeros(wl.spec.n, np.uint8)
    mask[800:] = 1
    net.set_output_mask(mask)
    seen_out = 0
    for t in range(80):
        net.begin_step(t)
        net.run_propagation()
        pend = net.download_pending()
        ring = net.ring()
        scale = np.float32(2.0 ** -net.fixed_shift)
        assert_bits_equal(pend, (ring[t % ring.shape[0]].astype(np.float32) * scale).astype(np.float32), "pending")
        out, full = net.neuron_step_outputs()
        assert np.array_equal(out, full[full >= 800])
        seen_out += out.size
        net.commit_firings()
    assert seen_out > 0


def test_reset_state_repeats_the_run(gpu):
    wl = W.populations((L.LIF_VJ, L.IZH), (1200, 300), k_in=60)
    net = wl.make_net()
    T = 150
    x0, l0 = net.get_state()
    net.record(0, T, counts=True, raster=True)
    net.run(T, L.ENGINE_GRAPH)
    r1, c1 = net.raster(), net.counts()
    xa, la = net.get_state()
    ev1 = net.events()
    net.reset_state()
    assert net.time == 0 and net.events() == 0
    x, last = net.get_state()
    # back to the state at finalize (the workload's initial state was uploaded after it)
    fresh = copy.copy(wl)
    fresh.x0 = None
    base = fresh.make_net()
    xb, lb = base.get_state()
    assert_bits_equal(x, xb, "state after reset")
    assert_bits_equal(last, lb, "LastSpike after reset")
    net.set_state(x0, l0)
    net.record(0, T, counts=True, raster=True)
    net.run(T, L.ENGINE_GRAPH)
    assert_bits_equal(net.raster(), r1, "raster of the repeated run")
    assert_bits_equal(net.counts(), c1, "counts of the repeated run")
    xr, lr = net.get_state()
    assert_bits_equal(xr, xa, "state of the repeated run")
    assert net.events() == ev1


def test_reset_state_keeps_weights_clears_traces(gpu):
    wl = W.stdp_izh(n=1500, k_in=60, steps=300, w_bg=2.0)
    net = wl.make_net()
    net.run(300, L.ENGINE_GRAPH)
    w = net.weights()
    assert not np.array_equal(w, wl.syn.weight)
    net.reset_state()
    assert_bits_equal(net.weights(), w, "weights kept")
    apre, apost, last = net.stdp_state()
    assert not apre.any() and not apost.any() and not last.any()
