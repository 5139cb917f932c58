#!/usr/bin/env python3
"""Closed-loop benchmark of the sparse LIF network on one MI355X.

Workload: 100,000 LIF voltage-jump neurons (80% excitatory), fixed in-degree 1,000
(1e8 synapses, CSR, excitatory +0.1 mV / inhibitory -0.5 mV), integer delays
U{1..20} in a delay ring, tau_m 20 ms, V_th -50, V_reset -65, refractory 2 ms,
RK4 with dt 0.1 ms, Poisson background of 1,000 sources x 10 Hz per neuron (0.1 mV).
A step is one simulated 0.1 ms tick of the whole network; inputs are resident in
HBM before the timed region. Prints one JSON line.

  value      neuron-steps/s (N x steps / time), with syn-events/s beside it
  roofline   the dominant kernel (the persistent step kernel) timed with HIP events
             on its own stream; algorithmic bytes = 24 B per neuron-step + 8 B per
             synaptic event. A second entry covers replay delivery (the recorded spikes of
             --replay-steps steps injected in one launch; 8 B per event + 4 B per output arrival), and
             a third replay STDP (the raster of --stdp-steps steps of the STDP network in one pass over
             its synapses; 36 B per plastic synapse).
  cpu_baseline  the CPU oracle in the legacy engine's queue-parallel structure (per-event float
             delivery), on every usable core for --cpu-steps steps and on one core for
             --cpu-steps-one steps of the same network, best of 3 after a warm-up, lscpu recorded
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
STATE_BYTES = 24   # per neuron-step: read V, refractory counter, ring slot; write V, counter; clear slot
EVENT_BYTES = 8    # per synaptic event: packed index/delay (4 B) + weight (4 B)


def measured_traffic(kernel: str, steps: int | None = None):
    """HBM bytes per launch of `kernel` from the committed PMC profile (profiles/traffic.json),
    scaled to `steps` steps when the launch length differs; None when not profiled."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        k = t[kernel]
    except (OSError, KeyError, ValueError):
        return None
    b = (k["fetch_kib"] * t["fetch_factor"] + k["write_kib"] * t["write_factor"]) * 1024.0
    if steps is not None:
        b *= steps / k["steps_per_launch"]
    return int(b)


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu() -> dict:
    """lscpu model, sockets, cores and threads of this host, and the CPUs this process may use."""
    info = {"model": None, "sockets": None, "cores_per_socket": None, "threads_per_core": None}
    keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
            "Thread(s) per core": "threads_per_core"}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                v = v.strip()
                info[keys[k.strip()]] = int(v) if v.isdigit() else v
    except Exception:
        pass
    info["usable_cpus"] = len(os.sched_getaffinity(0))
    return info


def cpu_baseline(wl, steps: int, steps_one: int) -> dict:
    """The CPU oracle in the legacy engine's structure (a thread per queue of target neurons,
    per-queue synapse lists, per-event float delivery, orc_run_queues): `steps` steps on every usable
    core and `steps_one` steps on one; each is the best of 3 runs from the initial state after one
    warm-up run; the stepping loop only is timed (no network build)."""
    import ctypes as C
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests"))
    import oracle as O
    from harness import OracleRun
    cpu = host_cpu()
    threads = max(1, min(cpu["usable_cpus"], int(os.environ.get("OMP_NUM_THREADS", "0") or cpu["usable_cpus"])))
    n = wl.spec.n

    def best_of(nq: int, nsteps: int):
        times, events = [], 0
        for rep in range(4):  # 1 warm-up + 3 timed
            run = OracleRun(wl)
            counts = np.zeros(nsteps, np.int32)
            secs = C.c_double()
            st = O.lib().orc_run_queues(C.byref(run.net), nq, O.ptr(run.x, C.c_float), O.ptr(run.last, C.c_int32),
                                        O.ptr(run.hist, C.c_int32), O.ptr(run.hist_cnt, C.c_int32), 0, nsteps,
                                        O.ptr(counts, C.c_int32), O.ptr(run.events, C.c_int64), C.byref(secs))
            assert st == 0, "orc_run_queues failed"
            if rep:
                times.append(secs.value)
            events = int(run.events[0])
        return min(times), events

    t_all, ev_all = best_of(threads, steps)
    t_one, ev_one = best_of(1, steps_one)
    return {"value": n * steps / t_all, "unit": "neuron-steps/s", "cores": threads, "kind": "port",
            "syn_events_per_s": ev_all / t_all,
            "sample": f"CPU oracle in the legacy engine's queue structure (one OpenMP thread per queue of target "
                      f"neurons, per-event float delivery), steps 0..{steps - 1} of the same {n}-neuron / "
                      f"{wl.syn.count}-synapse network on {threads} threads, best of 3 after a warm-up: {t_all:.2f} s",
            "single_thread": {"value": n * steps_one / t_one, "unit": "neuron-steps/s", "cores": 1,
                              "syn_events_per_s": ev_one / t_one,
                              "sample": f"steps 0..{steps_one - 1} on 1 thread, best of 3 after a warm-up: "
                                        f"{t_one:.2f} s"},
            "cpu": cpu}


def replay_roofline(net, wl, raster: np.ndarray, reps: int) -> dict:
    from neuronbridge_amd import _lib as L
    from neuronbridge_amd.hiptime import DeviceBuffer, Event
    K, n = raster.shape[0], wl.spec.n
    bits = np.unpackbits(raster.view(np.uint8), axis=1, bitorder="little")[:, :n]
    spikes = [np.nonzero(b)[0].astype(np.int32) for b in bits]
    tp = np.zeros(K + 1, np.int32)
    tp[1:] = np.cumsum([s.size for s in spikes])
    ids = np.concatenate(spikes).astype(np.int32) if tp[-1] else np.zeros(1, np.int32)
    d_ids, d_tp, d_out = DeviceBuffer(ids.nbytes), DeviceBuffer(tp.nbytes), DeviceBuffer(K * n * 4)
    d_ids.upload(ids)
    d_tp.upload(tp)
    deg = np.diff(wl.syn.row_ptr)
    events = int(deg[ids[:tp[-1]]].sum())
    stream = net.stream
    L.call("nbx_replay_deliver_device", net.h, d_ids.ptr, d_tp.ptr, K, d_out.ptr)  # warm
    net.sync()
    e0, e1 = Event(), Event()
    e0.record(stream)
    for _ in range(reps):
        L.call("nbx_replay_deliver_device", net.h, d_ids.ptr, d_tp.ptr, K, d_out.ptr)
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / reps
    alg = events * EVENT_BYTES + K * n * 4
    ach = alg / (ms * 1e-3) / 1e9
    kernel = {L.REPLAY_STREAM: "nbx_replay_stream", L.REPLAY_SLOTS: "nbx_replay_slots",
              L.REPLAY_OFFSETS: "nbx_replay_ring"}.get(net.replay_path(),
                                                                                           "nbx_replay_global")
    return {"kernel": kernel, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": measured_traffic(kernel, K),
            "algorithmic_bytes": alg,
            "steps_per_launch": K, "events_per_launch": events, "ms_per_launch": round(ms, 4),
            "syn_events_per_s": events / (ms * 1e-3)}


def batched_bench(steps: int, sample: int) -> dict:
    """Config 4: 256 independent LIF instances (N ~ U{512..8192}, dense BF16 weights) in one
    launch schedule vs the plain per-instance FP32 path on a sample of instances (graph engine,
    same BF16-valued weights); rasters of the sampled instances must be identical."""
    from neuronbridge_amd import _lib as L
    from neuronbridge_amd import workloads as W
    from neuronbridge_amd.batch import Batch
    from neuronbridge_amd.hiptime import Event
    t0 = time.perf_counter()
    inst, shared = W.batched_lif(steps=steps)
    log(f"batched: built {len(inst)} instances ({sum(i.n for i in inst)} neurons, "
        f"{sum(i.n * i.n for i in inst) / 1e9:.2f} G BF16 weights) in {time.perf_counter() - t0:.1f} s")
    b = Batch(inst, **shared)
    thr = b.tune(100)
    parts = b.parts()
    log(f"batched: tuned to {'split' if parts.any() else 'one workgroup per instance'}, {thr} threads"
        + (f", {int(parts.sum())} parts (max {int(parts.max())} per instance)" if parts.any() else ""))
    b.record(0, steps)
    e0, e1 = Event(), Event()
    e0.record(b.stream)
    b.run_async(steps)
    e1.record(b.stream)
    e1.synchronize()
    ms_b = e0.elapsed_ms(e1)
    ev_b = b.events()
    total_n = sum(i.n for i in inst)
    order = np.argsort([i.n for i in inst])
    pick = order[np.linspace(0, len(inst) - 1, sample).astype(int)] if sample > 0 else []
    plain_ms, plain_n, identical = 0.0, 0, True
    warm = 64  # the first 64 steps capture and instantiate the step graph; they are not timed
    for k in pick:
        ins = inst[int(k)]
        net = W.instance_workload(ins, shared).make_net(fixed_shift=ins.fixed_shift)
        net.record(0, steps, counts=False, raster=True)
        net.run(warm, L.ENGINE_GRAPH)
        f0, f1 = Event(), Event()
        f0.record(net.stream)
        net.run_async(steps - warm, L.ENGINE_GRAPH)
        f1.record(net.stream)
        f1.synchronize()
        net.sync()
        plain_ms += f0.elapsed_ms(f1)
        plain_n += ins.n
        identical &= bool(np.array_equal(net.raster(), b.raster(int(k))))
        net.close()
    batched_rate = total_n * steps / (ms_b * 1e-3)
    out = {"workload": f"{len(inst)} LIF instances, N~U{{512..8192}}, dense BF16 W, RK4 dt 0.1 ms, {steps} steps",
           "value": batched_rate, "unit": "instance-neuron-steps/s", "ms": round(ms_b, 2),
           "syn_events_per_s": ev_b / (ms_b * 1e-3), "workgroup_threads": thr,
           "schedule": ({"kind": "split", "parts": int(parts.sum()), "split_instances": int((parts > 1).sum()),
                         "max_parts": int(parts.max())} if parts.any() else {"kind": "one workgroup per instance"}),
           # BF16 weight-row bytes the gather reads (2 B per synaptic event) per second
           "weight_row_gbs": round(ev_b * 2 / (ms_b * 1e-3) / 1e9, 1)}
    if plain_n:
        plain_rate = plain_n * (steps - warm) / (plain_ms * 1e-3)
        out.update({"plain_per_instance": {"value": plain_rate, "unit": "instance-neuron-steps/s",
                                           "sample_instances": [int(inst[int(k)].n) for k in pick]},
                    "speedup_vs_plain": batched_rate / plain_rate, "rasters_identical_on_sample": identical})
    return out


def stdp_bench(steps: int) -> dict:
    """Config 3: STDP on current-based Izhikevich, 5e4 neurons, in-degree 500 (2.5e7 synapses, E->E plastic),
    delays 1..10, dt 0.5 ms. Plastic updates = pre-side LTD at each arrival on a plastic synapse + post-side LTP on
    each plastic input of a firing neuron, counted from the recorded raster; 12 B per plastic update (index +
    weight read, weight write; traces are per neuron in the reference layout, per synapse here)."""
    from neuronbridge_amd import _lib as L
    from neuronbridge_amd import workloads as W
    from neuronbridge_amd.hiptime import Event
    t0 = time.perf_counter()
    wl = W.stdp_izh(steps=steps + 200)
    log(f"stdp: built {wl.spec.n} neurons / {wl.syn.count} synapses in {time.perf_counter() - t0:.1f} s")
    net = wl.make_net()
    engine = net.pick_engine()
    net.run(200, engine)
    net.record(net.time, steps, counts=True, raster=True)
    ev0 = net.events()
    e0, e1 = Event(), Event()
    net.sync()
    e0.record(net.stream)
    net.run_async(steps, engine)
    e1.record(net.stream)
    e1.synchronize()
    net.sync()
    ms = e0.elapsed_ms(e1)
    n = wl.spec.n
    bits = np.unpackbits(net.raster().view(np.uint8), axis=1, bitorder="little")[:, :n]
    fired = bits.sum(axis=0).astype(np.int64)
    s = wl.syn
    pre = np.repeat(np.arange(n), np.diff(s.row_ptr))
    plastic = s.plastic.astype(bool)
    out_plastic = np.bincount(pre[plastic], minlength=n)
    in_plastic = np.bincount(s.post[plastic], minlength=n)
    updates = int((fired * out_plastic).sum() + (fired * in_plastic).sum())
    secs = ms * 1e-3
    w_after = net.weights()
    syn_events = net.events() - ev0
    net.close()
    return {"workload": f"STDP Izhikevich {n} neurons, in-degree 500, {int(plastic.sum())} plastic synapses, "
                        f"delays U{{1..10}}, dt 0.5 ms, {steps} steps",
            "engine": {L.ENGINE_GRAPH: "graph", L.ENGINE_KERNELS: "kernels"}.get(engine, str(engine)),
            "ms_per_step": ms / steps, "neuron_steps_per_s": n * steps / secs,
            "syn_events_per_s": syn_events / secs,
            "plastic_updates_per_s": updates / secs,
            "plastic_update_gbs": round(updates * 12 / secs / 1e9, 1),
            "spike_rate_hz": float(fired.sum()) / n / (steps * 0.5e-3),
            "weights_in_range": bool(np.all(w_after[plastic] >= 0.0) and np.all(w_after[plastic] <= 1.0))}


REPLAY_STDP_BYTES = 36  # per plastic synapse and launch: index+delay, weight, Apre/Apost/last in; weight, traces out


def replay_stdp_bench(steps: int, reps: int) -> dict:
    """Replay STDP on the STDP configuration's network: the raster of `steps` closed-loop steps (from
    step 0) replayed in one launch sequence (spike lists per neuron, then one pass over the synapses).
    The first replay must leave the closed loop's weights and trace records bit for bit; the timed
    launches replay the same raster at later steps (same work). Algorithmic bytes: 36 B per plastic
    synapse per launch (each plastic synapse's state read and written once, whatever the step count)."""
    from neuronbridge_amd import _lib as L
    from neuronbridge_amd import workloads as W
    from neuronbridge_amd.hiptime import DeviceBuffer, Event
    wl = W.stdp_izh(steps=steps)
    n = wl.spec.n
    ref = wl.make_net()
    engine = ref.pick_engine()
    ref.record(0, steps, counts=True, raster=True)
    e0, e1 = Event(), Event()
    ref.sync()
    e0.record(ref.stream)
    ref.run_async(steps, engine)
    e1.record(ref.stream)
    e1.synchronize()
    ref.sync()
    loop_ms = e0.elapsed_ms(e1)
    bits = np.unpackbits(ref.raster().view(np.uint8), axis=1, bitorder="little")[:, :n]
    w_ref = ref.weights()
    tr_ref = ref.stdp_state()
    ref.close()
    spikes = [np.nonzero(b)[0].astype(np.int32) for b in bits]
    tp = np.zeros(steps + 1, np.int32)
    tp[1:] = np.cumsum([x.size for x in spikes])
    ids = np.concatenate(spikes).astype(np.int32) if tp[-1] else np.zeros(1, np.int32)
    d_ids, d_tp, d_upd = DeviceBuffer(ids.nbytes), DeviceBuffer(tp.nbytes), DeviceBuffer(8)
    d_ids.upload(ids)
    d_tp.upload(tp)
    d_upd.upload(np.zeros(1, np.uint64))
    net = wl.make_net()
    L.call("nbx_replay_stdp_device", net.h, d_ids.ptr, d_tp.ptr, 0, 0, steps, int(tp[-1]), d_upd.ptr)
    net.sync()
    upd = np.zeros(1, np.uint64)
    d_upd.download(upd)
    updates = int(upd[0])
    same = np.array_equal(net.weights().view(np.uint32), w_ref.view(np.uint32))
    for a, b in zip(net.stdp_state(), tr_ref):
        same = same and np.array_equal(a.view(np.uint32) if a.dtype == np.float32 else a,
                                       b.view(np.uint32) if b.dtype == np.float32 else b)
    e0.record(net.stream)
    for r in range(reps):
        L.call("nbx_replay_stdp_device", net.h, d_ids.ptr, d_tp.ptr, (r + 1) * steps, 0, steps, int(tp[-1]), None)
    e1.record(net.stream)
    e1.synchronize()
    net.sync()
    ms = e0.elapsed_ms(e1) / reps
    plastic = int(wl.syn.plastic.sum())
    net.close()
    alg = plastic * REPLAY_STDP_BYTES
    ach = alg / (ms * 1e-3) / 1e9
    return {"workload": f"replay STDP: raster of {steps} closed-loop steps of the STDP network ({n} Izhikevich "
                        f"neurons, in-degree 500, {plastic} plastic synapses, delays U{{1..10}}, dt 0.5 ms) "
                        f"in one launch sequence",
            "ms_per_launch": round(ms, 4), "plastic_updates_per_launch": updates,
            "plastic_updates_per_s": updates / (ms * 1e-3),
            "closed_loop_ms_same_steps": round(loop_ms, 3), "speedup_vs_closed_loop": round(loop_ms / ms, 1),
            "identical_to_closed_loop": bool(same),
            "roofline": {"kernel": "nbx_replay_stdp", "bound": "hbm", "achieved": round(ach, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                         "traffic": measured_traffic("nbx_replay_stdp"), "algorithmic_bytes": alg,
                         "timed": "whole launch sequence (spike lists + synapse pass)"}}


def dense_bench(steps: int, noise_scale: float = 1.0) -> dict:
    """Config 1: dense all-to-all current-based Izhikevich, 1,000 neurons, dt 0.1 ms, Gaussian input currents
    (noise_scale multiplies them; sqrt(10) restores the per-millisecond variance of the 1 ms original)."""
    from neuronbridge_amd import _lib as L
    from neuronbridge_amd import workloads as W
    from neuronbridge_amd.hiptime import Event
    wl = W.dense_izh(steps=steps + 100, noise_scale=noise_scale)
    names = {L.ENGINE_GRAPH: "graph", L.ENGINE_KERNELS: "kernels", L.ENGINE_PERSISTENT: "persistent"}

    def timed(engine):
        net = wl.make_net()
        net.run(100, engine)
        net.record(net.time, steps, counts=True, raster=True)
        ev0 = net.events()
        e0, e1 = Event(), Event()
        net.sync()
        e0.record(net.stream)
        net.run_async(steps, engine)
        e1.record(net.stream)
        e1.synchronize()
        net.sync()
        out = (e0.elapsed_ms(e1), net.events() - ev0, int(net.counts().sum()), net.raster())
        net.close()
        return out

    engine = wl.make_net().pick_engine()
    ms, ev, spikes, raster = timed(engine)
    res = {"workload": f"dense Izhikevich {wl.spec.n} neurons all-to-all, Gaussian currents (std 5 exc / 2 inh "
                       f"per step x {noise_scale:.3g}), dt 0.1 ms, {steps} steps",
           "engine": names.get(engine, str(engine)),
           "ms_per_step": ms / steps, "neuron_steps_per_s": wl.spec.n * steps / (ms * 1e-3),
           "syn_events_per_s": ev / (ms * 1e-3),
           "spike_rate_hz": spikes / wl.spec.n / (steps * 0.1e-3)}
    if engine != L.ENGINE_GRAPH:  # the graph engine on the same steps, and the same raster
        gms, _, _, graster = timed(L.ENGINE_GRAPH)
        res["graph_ms_per_step"] = gms / steps
        res["speedup_vs_graph"] = round(gms / ms, 2)
        res["raster_identical_to_graph"] = bool(np.array_equal(raster, graster))
    return res


def models_bench(steps: int) -> dict:
    """The remaining legacy models on the config-2 topology (1e5 neurons, in-degree 1000, delays 1..20):
    conductance LIF (double / triple exponential, RK4), PoissonRate and the trigger relay, each on the
    engine the library picks; closed-loop stepping after 100 warm steps."""
    from neuronbridge_amd import _lib as L
    from neuronbridge_amd import workloads as W
    from neuronbridge_amd.hiptime import Event
    from neuronbridge_amd.network import build_fixed_indegree
    n, k_in, seed = 100_000, 1000, W.SEED_BASE + 6
    t0 = time.perf_counter()
    topo = build_fixed_indegree(n, k_in, 1, 20, seed, int(round(0.8 * n)))
    log(f"models: topology built in {time.perf_counter() - t0:.1f} s")
    names = {L.LIF_EXP2: "lif_exp_double", L.LIF_EXP3: "lif_exp_triple", L.POISSON: "poisson_rate",
             L.TRIGGER_RELAY: "trigger_relay"}
    out = {"workload": f"config-2 topology (1e5 neurons, in-degree 1000, delays U{{1..20}}), dt 0.1 ms, "
                       f"{steps} steps per model"}
    for model, name in names.items():
        integ = L.RK4 if model in (L.LIF_EXP2, L.LIF_EXP3) else L.EULER
        wl = W.model_sparse(model, n=n, k_in=k_in, integrator=integ, seed=seed, topology=topo)
        net = wl.make_net()
        engine = net.pick_engine()
        net.run(100, engine)
        net.record(net.time, steps, counts=True, raster=False)
        ev0 = net.events()
        e0, e1 = Event(), Event()
        net.sync()
        e0.record(net.stream)
        net.run_async(steps, engine)
        e1.record(net.stream)
        e1.synchronize()
        net.sync()
        ms = e0.elapsed_ms(e1)
        ev = net.events() - ev0
        spikes = int(net.counts().sum())
        net.close()
        out[name] = {"engine": {L.ENGINE_GRAPH: "graph", L.ENGINE_KERNELS: "kernels",
                                L.ENGINE_PERSISTENT: "persistent"}.get(engine, str(engine)),
                     "ms_per_step": ms / steps, "neuron_steps_per_s": n * steps / (ms * 1e-3),
                     "syn_events_per_s": ev / (ms * 1e-3), "spike_rate_hz": spikes / n / (steps * 0.1e-3)}
        log(f"models: {name}: {out[name]}")
    return out


def rules_bench(steps: int) -> dict:
    """Trigger-driven learning rules at config-2 size: 1e5 LIF neurons, in-degree 1000, 80% of E->E
    synapses split over R-STDP, additive-kernel and cerebellar rules, 3% of all synapses triggers
    (graph engine)."""
    from neuronbridge_amd import _lib as L
    from neuronbridge_amd import workloads as W
    from neuronbridge_amd.hiptime import Event
    wl = W.trigger_rules(n=100_000, k_in=1000, cerebellar=True)
    net = wl.make_net()
    engine = net.pick_engine()
    w0 = net.weights()
    net.run(100, engine)
    net.record(net.time, steps, counts=True, raster=False)
    ev0 = net.events()
    e0, e1 = Event(), Event()
    net.sync()
    e0.record(net.stream)
    net.run_async(steps, engine)
    e1.record(net.stream)
    e1.synchronize()
    net.sync()
    ms = e0.elapsed_ms(e1)
    ev = net.events() - ev0
    spikes = int(net.counts().sum())
    rule = wl.rules["rule"]
    changed = net.weights() != w0
    net.close()
    n = wl.spec.n
    return {"workload": "config-2 LIF network (1e5 neurons, 1e8 synapses) with R-STDP, additive-kernel and "
                        f"cerebellar rules on 80% of E->E synapses, 3% triggers, {steps} steps",
            "engine": {L.ENGINE_GRAPH: "graph", L.ENGINE_KERNELS: "kernels"}.get(engine, str(engine)),
            "ms_per_step": ms / steps, "neuron_steps_per_s": n * steps / (ms * 1e-3),
            "syn_events_per_s": ev / (ms * 1e-3), "spike_rate_hz": spikes / n / (steps * 0.1e-3),
            "rstdp_weights_changed": float(changed[rule == 0].mean()),
            "additive_weights_changed": float(changed[rule == 1].mean()),
            "cerebellar_weights_changed": float(changed[rule == 2].mean())}


def adjoint_bench(reps: int) -> dict:
    """Config 5: forward + surrogate-gradient backward (checkpoint 50) of the 2048-neuron recurrent
    LIF network, 64 trials x 500 steps; GEMM flop count over the timed training step."""
    import torch
    from neuronbridge_amd import workloads as W
    from neuronbridge_amd.adjoint import LifParams, autograd_loss_and_grads, loss_and_grads
    Wi, Wr, Wo, x, tg = W.diff_lif()
    dev = "cuda:0"
    args = [torch.tensor(a, device=dev) for a in (Wi, Wr, Wo, x, tg)]
    p = LifParams()
    loss_and_grads(*args, p, 50)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        loss, grads, S = loss_and_grads(*args, p, 50)
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / reps
    T, cin, B = x.shape
    N, M = Wr.shape[0], Wo.shape[0]
    rec = 2.0 * N * N * B * T
    flop = 4 * rec + 2 * 2.0 * N * cin * B * T + 3 * 2.0 * M * N * B * T
    # the plain path: the same network unrolled with torch ops and differentiated by autograd (FP32,
    # every step's tensors held), on the adjoint's spike train so the gradients are comparable
    autograd_loss_and_grads(*args, p, spikes=S)
    torch.cuda.synchronize()
    ev0.record()
    loss_p, grads_p, _ = autograd_loss_and_grads(*args, p, spikes=S)
    ev1.record()
    torch.cuda.synchronize()
    ms_p = ev0.elapsed_time(ev1)
    rel = max(float((a - b).norm() / b.norm()) for a, b in zip(grads, grads_p))
    return {"workload": f"recurrent LIF N={N}, {B} trials, T={T}, dt 1 ms, 200 Poisson inputs, 8 readouts, "
                        f"checkpoint 50", "ms_per_training_step": round(ms, 3),
            "gemm_tflops": flop / (ms * 1e-3) / 1e12, "gemm_peak_tflops_fp32": 157.3,
            "loss": float(loss), "spike_rate_hz": float(S.float().mean()) * 1000.0,
            "plain_autograd": {"ms_per_training_step": round(ms_p, 3),
                               "what": "torch autograd through the unrolled FP32 step on the GPU, same spikes"},
            "speedup_vs_plain": ms_p / ms, "grad_rel_err_vs_plain": rel, "grad_tol": 1e-4}


def adjoint_kernels_bench(steps: int) -> dict:
    """The C5 step kernels alone, without torch (device buffers through HIP directly, so a counter
    profile holds libnbx kernels only): one forward segment of `steps` steps (binary spikes, the
    sum of the W_rec^T rows of the neurons that fired) and one backward segment (W_rec^T dI_{t+1}
    on FP32 MFMA + the adjoint step) at the configuration's shape, N = 2048, 64 trials, on the
    configuration's weights and inputs."""
    import ctypes as C
    from neuronbridge_amd import _lib as L
    from neuronbridge_amd import workloads as W
    from neuronbridge_amd.adjoint import LifParams
    from neuronbridge_amd.hiptime import DeviceBuffer, Event, hip
    Wi, Wr, Wo, x, tg = W.diff_lif()
    N, B = Wr.shape[0], x.shape[2]
    p = LifParams()
    cp = p.c_params(N, B)
    nb = N * B
    IN = np.matmul(Wi, x[:steps]).astype(np.float32)  # [steps, N, B]
    d_WT = DeviceBuffer(N * N * 4)
    d_WT.upload(np.ascontiguousarray(Wr.T).astype(np.float32))
    d_W = DeviceBuffer(N * N * 4)
    d_W.upload(np.ascontiguousarray(Wr).astype(np.float32))
    d_IN = DeviceBuffer(IN.nbytes)
    d_IN.upload(np.ascontiguousarray(IN))
    d_V, d_s0 = DeviceBuffer(nb * 4), DeviceBuffer(nb * 4)
    d_s0.upload(np.zeros(nb, np.float32))
    d_S, d_Vm, d_dI = DeviceBuffer(steps * nb * 4), DeviceBuffer(steps * nb * 4), DeviceBuffer(steps * nb * 4)
    d_bits = DeviceBuffer(2 * B * ((N + 31) // 32) * 4)
    rng = np.random.default_rng(5)
    d_R, d_dV, d_dIa = DeviceBuffer(steps * nb * 4), DeviceBuffer(nb * 4), DeviceBuffer(nb * 4)
    d_R.upload((rng.standard_normal(steps * nb) * 1e-3).astype(np.float32))
    d_dIa.upload((rng.standard_normal(nb) * 1e-3).astype(np.float32))
    def segments():
        d_V.upload(np.full(nb, p.v_rest, np.float32))
        d_dV.upload(np.zeros(nb, np.float32))
        hip().hipDeviceSynchronize()
        e0.record(0)
        L.call("nbx_lif_fwd_segment_sparse", C.byref(cp), d_WT.ptr, d_IN.ptr, d_s0.ptr, d_V.ptr, d_S.ptr, d_Vm.ptr,
               d_bits.ptr, steps, C.c_void_p(0))
        e1.record(0)
        L.call("nbx_lif_bwd_segment", C.byref(cp), d_WT.ptr, d_dIa.ptr, d_R.ptr, d_Vm.ptr, d_dV.ptr, d_dI.ptr, steps,
               C.c_void_p(0))
        e2.record(0)
        e2.synchronize()

    e0, e1, e2 = Event(), Event(), Event()
    segments()  # warm-up (first launches load the code objects)
    segments()
    fwd, bwd = e0.elapsed_ms(e1) / steps, e1.elapsed_ms(e2) / steps
    flop = 2.0 * N * N * B
    S = np.zeros(steps * nb, np.float32)
    d_S.download(S)
    return {"workload": f"C5 step kernels alone: N={N}, {B} trials, {steps}-step segments, no torch",
            "fwd_sparse_us_per_step": round(fwd * 1e3, 2), "bwd_mfma_us_per_step": round(bwd * 1e3, 2),
            "bwd_tflops": flop / (bwd * 1e-3) / 1e12, "spike_rate_hz": float(S.mean()) * 1000.0}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--engine", default="auto", choices=["auto", "persistent", "graph", "kernels"])
    ap.add_argument("--integrator", default="rk4", choices=["rk4", "euler"])
    ap.add_argument("--cpu-steps", type=int, default=10000, help="CPU baseline steps on every usable core")
    ap.add_argument("--cpu-steps-one", type=int, default=1000, help="CPU baseline steps on one core")
    ap.add_argument("--replay-reps", type=int, default=5)
    ap.add_argument("--replay-steps", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-replay", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=5000)
    ap.add_argument("--batched-sample", type=int, default=8)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-kernels", type=int, default=0,
                    help="also time the C5 step kernels alone over segments of this many steps (no torch)")
    ap.add_argument("--no-stdp", action="store_true")
    ap.add_argument("--stdp-steps", type=int, default=2000)
    ap.add_argument("--no-dense", action="store_true")
    ap.add_argument("--no-models", action="store_true")
    ap.add_argument("--models-steps", type=int, default=1000)
    ap.add_argument("--no-rules", action="store_true")
    ap.add_argument("--rules-steps", type=int, default=500)
    ap.add_argument("--dense-steps", type=int, default=10000)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="also run one device-to-device copy of this size (PMC byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (this path runs on a single "
              "MI355X; multi-GPU sharding is described in the design notes, not benchmarked)", file=sys.stderr)
        return 2
    import neuronbridge_amd as nb
    from neuronbridge_amd import _lib as L
    from neuronbridge_amd import workloads as W
    from neuronbridge_amd.hiptime import Event

    integ = nb.RK4 if args.integrator == "rk4" else nb.EULER
    t0 = time.perf_counter()
    wl = W.sparse_lif(integrator=integ, steps=args.warmup + args.steps)
    log(f"built {wl.spec.n} neurons / {wl.syn.count} synapses in {time.perf_counter() - t0:.1f} s")
    net = wl.make_net()
    engines = {"auto": L.ENGINE_AUTO, "persistent": L.ENGINE_PERSISTENT, "graph": L.ENGINE_GRAPH,
               "kernels": L.ENGINE_KERNELS}
    engine = engines[args.engine]
    if engine == L.ENGINE_AUTO:
        engine = net.pick_engine()
    ename = {v: k for k, v in engines.items()}[engine]
    log(f"engine {ename}, fixed-point shift {net.fixed_shift}")
    net.run(args.warmup, engine)
    ev0 = net.events()
    log(f"warmup {args.warmup} steps done")
    stream = net.stream
    e0, e1 = Event(), Event()
    net.sync()
    w0 = time.perf_counter()
    e0.record(stream)
    net.run_async(args.steps, engine)
    e1.record(stream)
    e1.synchronize()
    wall = time.perf_counter() - w0
    net.sync()
    ms = e0.elapsed_ms(e1)
    events = net.events() - ev0
    n = wl.spec.n
    secs = ms * 1e-3
    nsteps_s = n * args.steps / secs
    log(f"timed {args.steps} steps: {ms:.2f} ms (wall {wall * 1e3:.2f} ms), {events} events")
    alg = STATE_BYTES * n * args.steps + EVENT_BYTES * events
    ach = alg / secs / 1e9
    kname = "nbx_persistent" if engine == L.ENGINE_PERSISTENT else "step graph"
    out = {
        "metric": "neuron-steps/s & syn-events/s, 1 MI355X; batched+adjoint vs plain at tol",
        "value": nsteps_s, "unit": "neuron-steps/s",
        "syn_events_per_s": events / secs,
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded fixed in-degree topology, Philox Poisson background)",
        "config": {"workload": f"sparse LIF 1e5 neurons, in-degree 1000, delays U{{1..20}}, "
                               f"{args.integrator} dt 0.1 ms", "neurons": n, "synapses": int(wl.syn.count),
                   "engine": ename, "fixed_shift": net.fixed_shift, "parallelism": "single GPU"},
        "roofline": {"kernel": kname, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic("nbx_persistent", args.steps) if kname == "nbx_persistent" else None,
                     "launch_ms": round(ms, 3), "algorithmic_bytes": alg},
    }
    if not args.no_replay:
        K = args.replay_steps
        t_rec = net.time
        net.record(t_rec, K, counts=True, raster=True)
        net.run(K, engine)
        out["replay_roofline"] = replay_roofline(net, wl, net.raster(), args.replay_reps)
        log(f"replay: {out['replay_roofline']}")
    if args.calibrate_copy_mib > 0:
        from neuronbridge_amd.hiptime import DeviceBuffer, hip
        nbytes = args.calibrate_copy_mib << 20
        src, dst = DeviceBuffer(nbytes), DeviceBuffer(nbytes)
        rc = hip().hipMemcpy(dst.ptr, src.ptr, nbytes, 3)  # device to device
        hip().hipDeviceSynchronize()
        log(f"calibration copy of {args.calibrate_copy_mib} MiB (rc {rc})")
    if not args.no_batched:
        net.close()
        out["batched"] = batched_bench(args.batched_steps, args.batched_sample)
        log(f"batched: {out['batched']}")
    if not args.no_adjoint:
        out["adjoint"] = adjoint_bench(3)
        log(f"adjoint: {out['adjoint']}")
    if args.adjoint_kernels > 0:
        out["adjoint_kernels"] = adjoint_kernels_bench(args.adjoint_kernels)
        log(f"adjoint kernels: {out['adjoint_kernels']}")
    if not args.no_stdp:
        out["stdp"] = stdp_bench(args.stdp_steps)
        log(f"stdp: {out['stdp']}")
        out["replay_stdp"] = replay_stdp_bench(args.stdp_steps, args.replay_reps)
        log(f"replay stdp: {out['replay_stdp']}")
    if not args.no_dense:
        out["dense"] = dense_bench(args.dense_steps)
        log(f"dense: {out['dense']}")
        out["dense_active"] = dense_bench(args.dense_steps, noise_scale=10 ** 0.5)
        log(f"dense (active): {out['dense_active']}")
    if not args.no_models:
        out["models"] = models_bench(args.models_steps)
    if not args.no_rules:
        out["rules"] = rules_bench(args.rules_steps)
        log(f"rules: {out['rules']}")
    if not args.no_cpu and args.cpu_steps > 0:
        log(f"CPU oracle baseline over {args.cpu_steps} steps ...")
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_steps, max(1, args.cpu_steps_one))
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
