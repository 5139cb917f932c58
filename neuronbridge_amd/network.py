"""Host-side wrapper of one simulated network on the GPU (numpy in, numpy out).

Mirrors the reference runtime's per-step interface on this path:
``GpuPropagationRuntime`` (BeginStep / RunPropagationOnly /
CommitCurrentDeviceFiringsFromExternal / ApplyPostLearningFromDeviceFirings,
src/native/gpu/include/gpu_runtime/GpuPropagationRuntime.h:85-144) and
``DenseUnifiedNeuronRuntime::Step`` (DenseUnifiedNeuronRuntime.h:28-33), plus
whole-run stepping (``Simulation.run``, python/src/neuronbridge/__init__.py:191-193).
Every method goes through libnbx.so; nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L


def _p(a, ctype):
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(ctype))


@dataclass
class ModelSpec:
    """Neuron population parameters (reference model defaults where the model has them)."""
    model: int
    n: int
    dt: float
    integrator: int = L.EULER
    t_ref: int = 0
    # LIF (TimeDrivenLIF_Voltage_jump.h:22-32 / TimeDrivenLIF_Exponential_Decay.h defaults differ;
    # callers state their own)
    v_rest: float = -65.0
    tau: float = 20.0
    v_th: float = -50.0
    r: float = 1.0
    v_reset: float = -65.0
    g_tau: float = 12.0
    e_rev: float = 0.0
    # Izhikevich (TimeDrivenIzhikevic_Exponential_Decay.h:28-52)
    izh_v_th: float = 30.0
    izh_r: float = 1.0
    e_ampa: float = 0.0
    e_gaba: float = -80.0
    ampa_tau: float = 5.0
    gaba_tau: float = 10.0
    # triple-exponential LIF NMDA conductance (TimeDrivenLIF_Exponential_triple.h:17)
    nmda_tau: float = 10.0
    # PoissonRate (PoissonRate.h:12-13)
    rate_bias_hz: float = 0.0
    rate_gain_hz: float = 1.0
    a: np.ndarray | None = None
    b: np.ndarray | None = None
    c: np.ndarray | None = None
    d: np.ndarray | None = None
    _keep: list = field(default_factory=list, repr=False)

    def desc(self) -> L.ModelDesc:
        arrs = []
        for v in (self.a, self.b, self.c, self.d):
            arrs.append(None if v is None else np.ascontiguousarray(v, dtype=np.float32))
        self._keep = arrs
        return L.ModelDesc(
            model=self.model, integrator=self.integrator, n=self.n, t_ref=self.t_ref, dt=self.dt,
            v_rest=self.v_rest, tau=self.tau, v_th=self.v_th, r=self.r, v_reset=self.v_reset,
            g_tau=self.g_tau, e_rev=self.e_rev, izh_v_th=self.izh_v_th, izh_r=self.izh_r,
            e_ampa=self.e_ampa, e_gaba=self.e_gaba, ampa_tau=self.ampa_tau, gaba_tau=self.gaba_tau,
            izh_a=_p(arrs[0], C.c_float), izh_b=_p(arrs[1], C.c_float),
            izh_c=_p(arrs[2], C.c_float), izh_d=_p(arrs[3], C.c_float),
            nmda_tau=self.nmda_tau, rate_bias_hz=self.rate_bias_hz, rate_gain_hz=self.rate_gain_hz)

    @property
    def nstate(self) -> int:
        return L.model_nstate(self.model)

    @property
    def channels(self) -> int:
        return L.model_channels(self.model)


@dataclass
class PopulationsSpec:
    """Several neuron populations in one network (nbx_net_create_populations): population k owns
    the next pops[k].n neurons; every population steps with the same dt. Arrivals are kept by
    synapse type, so the delay ring has the three typed channels."""
    pops: list

    @property
    def n(self) -> int:
        return int(sum(p.n for p in self.pops))

    @property
    def nstate(self) -> int:
        return max(p.nstate for p in self.pops)

    @property
    def channels(self) -> int:
        return 3

    @property
    def dt(self) -> float:
        return self.pops[0].dt

    @property
    def begins(self) -> list:
        out, b = [], 0
        for p in self.pops:
            out.append(b)
            b += p.n
        return out


@dataclass
class Synapses:
    """CSR by presynaptic neuron, rows sorted by (delay, post)."""
    row_ptr: np.ndarray  # int64 [n+1]
    post: np.ndarray     # int32 [S]
    delay: np.ndarray    # uint8 [S]
    type: np.ndarray     # uint8 [S]
    weight: np.ndarray   # float32 [S]
    max_delay: int
    plastic: np.ndarray | None = None  # uint8 [S]

    @property
    def count(self) -> int:
        return int(self.post.shape[0])


def build_fixed_indegree(n: int, k_in: int, d_min: int, d_max: int, seed: int, n_exc: int,
                         threads: int = 0) -> Synapses:
    """Random fixed in-degree topology built by libnbx's host builder (weights left 0)."""
    S = n * k_in
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    post = np.empty(S, dtype=np.int32)
    delay = np.empty(S, dtype=np.uint8)
    typ = np.empty(S, dtype=np.uint8)
    L.call("nbx_build_fixed_indegree", n, k_in, d_min, d_max, seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF,
           n_exc, threads, _p(row_ptr, C.c_int64), _p(post, C.c_int32), _p(delay, C.c_uint8), _p(typ, C.c_uint8))
    return Synapses(row_ptr, post, delay, typ, np.zeros(S, dtype=np.float32), d_max)


class Net:
    """One network resident on a GPU."""

    def __init__(self, spec, device: int = 0):
        """spec: a ModelSpec (one population) or a PopulationsSpec (several)."""
        self.spec = spec
        h = C.c_void_p()
        if isinstance(spec, PopulationsSpec):
            descs = (L.ModelDesc * len(spec.pops))(*[p.desc() for p in spec.pops])
            L.call("nbx_net_create_populations", descs, len(spec.pops), device, C.byref(h))
        else:
            desc = spec.desc()
            L.call("nbx_net_create", C.byref(desc), device, C.byref(h))
        self.h = h
        self.n = spec.n
        self.nstate = spec.nstate
        self.syn: Synapses | None = None
        self.syn_count = 0
        self._rec = (0, 0, 0)

    # ------------------------------------------------------------ configuration
    def set_sparse(self, syn: Synapses) -> None:
        self.syn = syn
        self.syn_count = syn.count
        d = L.SparseDesc(count=syn.count, max_delay=int(syn.max_delay),
                         row_ptr=_p(syn.row_ptr, C.c_int64), post=_p(syn.post, C.c_int32),
                         delay=_p(syn.delay, C.c_uint8), type=_p(syn.type, C.c_uint8),
                         weight=_p(syn.weight, C.c_float),
                         plastic=_p(syn.plastic, C.c_uint8) if syn.plastic is not None else None)
        L.call("nbx_net_set_sparse", self.h, C.byref(d))

    def load_nbnet(self, path: str, verify_checksum: bool = True) -> int:
        """Sparse synapses from a .nbnet connection table (sorted on the device); returns the count."""
        L.call("nbx_net_load_nbnet", self.h, os.fsencode(path), int(bool(verify_checksum)))
        self.syn_count = nbnet_info(path, False).connection_count
        return self.syn_count

    def set_dense(self, wt: np.ndarray, delay: int = 1) -> None:
        wt = np.ascontiguousarray(wt, dtype=np.float32)
        assert wt.shape == (self.n, self.n)
        L.call("nbx_net_set_dense", self.h, _p(wt, C.c_float), int(delay))

    def set_poisson(self, seed: int, lam: float, w_bg: float) -> None:
        L.call("nbx_net_set_poisson", self.h, seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF, float(lam), float(w_bg))

    def set_stdp(self, max_ltp, ltp_tau, max_ltd, ltd_tau, w_max) -> None:
        d = L.StdpDesc(max_ltp=max_ltp, ltp_tau=ltp_tau, max_ltd=max_ltd, ltd_tau=ltd_tau, w_max=w_max)
        L.call("nbx_net_set_stdp", self.h, C.byref(d))

    def set_rules(self, rules: list, rule: np.ndarray, trigger: np.ndarray, max_weight: np.ndarray) -> None:
        """Trigger-driven learning rules: `rules` is a list of dicts with kind (L.RULE_RSTDP /
        L.RULE_ADDITIVE) and its parameters; rule / trigger / max_weight are per synapse (CSR order)."""
        arr = (L.RuleDesc * max(1, len(rules)))()
        for k, r in enumerate(rules):
            for f, _ in L.RuleDesc._fields_:
                if f in r:
                    setattr(arr[k], f, r[f])
        self._rule_arrays = [np.ascontiguousarray(rule, dtype=np.int32), np.ascontiguousarray(trigger, dtype=np.int32),
                             np.ascontiguousarray(max_weight, dtype=np.float32)]
        ru, tr, mw = self._rule_arrays
        L.call("nbx_net_set_rules", self.h, arr, len(rules), _p(ru, C.c_int32), _p(tr, C.c_int32), _p(mw, C.c_float))

    def assign_nbnet_rules(self, rules: list) -> None:
        """Rules of a .nbnet-loaded network by the file's rule index (dicts as in set_rules; kind
        L.RULE_STDP maps onto pair STDP with per-record weight bounds)."""
        arr = (L.RuleDesc * max(1, len(rules)))()
        for k, r in enumerate(rules):
            for f, _ in L.RuleDesc._fields_:
                if f in r:
                    setattr(arr[k], f, r[f])
        L.call("nbx_net_assign_nbnet_rules", self.h, arr, len(rules))

    def rule_state(self):
        """(Apre, Apost, eligibility, last update) of the rule-plastic synapses (0 elsewhere)."""
        S = self.syn_count
        apre, apost, el = (np.zeros(S, np.float32) for _ in range(3))
        last = np.zeros(S, np.int32)
        L.call("nbx_net_get_rule_state", self.h, _p(apre, C.c_float), _p(apost, C.c_float), _p(el, C.c_float),
               _p(last, C.c_int32), S)
        return apre, apost, el, last

    def set_current(self, cur: np.ndarray, t_begin: int = 0) -> None:
        cur = np.ascontiguousarray(cur, dtype=np.float32)
        assert cur.ndim == 2 and cur.shape[1] == self.n
        L.call("nbx_net_set_current", self.h, _p(cur, C.c_float), int(t_begin), int(cur.shape[0]))

    def set_fixed_shift(self, shift: int) -> None:
        L.call("nbx_net_set_fixed_shift", self.h, int(shift))

    def finalize(self) -> "Net":
        L.call("nbx_net_finalize", self.h)
        return self

    @property
    def fixed_shift(self) -> int:
        s = C.c_int32()
        L.call("nbx_net_get_fixed_shift", self.h, C.byref(s))
        return s.value

    # ------------------------------------------------------------ state
    def set_state(self, x: np.ndarray, last: np.ndarray) -> None:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(self.n, self.nstate)
        last = np.ascontiguousarray(last, dtype=np.int32)
        L.call("nbx_net_set_state", self.h, _p(x, C.c_float), _p(last, C.c_int32))

    def get_state(self) -> tuple[np.ndarray, np.ndarray]:
        x = np.zeros((self.n, self.nstate), dtype=np.float32)
        last = np.zeros(self.n, dtype=np.int32)
        L.call("nbx_net_get_state", self.h, _p(x, C.c_float), _p(last, C.c_int32))
        return x, last

    # ------------------------------------------------------------ running
    def record(self, t_begin: int, nsteps: int, counts=True, raster=False) -> None:
        flags = (L.RECORD_COUNTS if counts else 0) | (L.RECORD_RASTER if raster else 0)
        L.call("nbx_net_record", self.h, int(t_begin), int(nsteps), flags)
        self._rec = (t_begin, nsteps, flags)

    def counts(self) -> np.ndarray:
        out = np.zeros(self._rec[1], dtype=np.int32)
        L.call("nbx_net_get_counts", self.h, _p(out, C.c_int32), out.shape[0])
        return out

    def raster(self) -> np.ndarray:
        words = (self.n + 31) // 32
        out = np.zeros((self._rec[1], words), dtype=np.uint32)
        L.call("nbx_net_get_raster", self.h, _p(out, C.c_uint32), out.size)
        return out

    def run(self, nsteps: int, engine: int = L.ENGINE_AUTO) -> None:
        L.call("nbx_net_run", self.h, int(nsteps), int(engine))

    def run_async(self, nsteps: int, engine: int = L.ENGINE_AUTO) -> None:
        L.call("nbx_net_run_async", self.h, int(nsteps), int(engine))

    def sync(self) -> None:
        L.call("nbx_net_sync", self.h)

    def pick_engine(self) -> int:
        e = C.c_int32()
        L.call("nbx_net_pick_engine", self.h, C.byref(e))
        return e.value

    @property
    def time(self) -> int:
        t = C.c_int32()
        L.call("nbx_net_time", self.h, C.byref(t))
        return t.value

    def events(self) -> int:
        e = C.c_int64()
        L.call("nbx_net_events", self.h, C.byref(e))
        return e.value

    @property
    def stream(self) -> int:
        return int(L.lib().nbx_net_stream(self.h) or 0)

    # ------------------------------------------------------------ per-phase calls
    def begin_step(self, t: int) -> None:
        L.call("nbx_net_begin_step", self.h, int(t))

    def run_propagation(self) -> None:
        L.call("nbx_net_run_propagation", self.h)

    def neuron_step(self) -> np.ndarray:
        ids = np.zeros(self.n, dtype=np.int32)
        cnt = C.c_int32()
        L.call("nbx_net_neuron_step", self.h, _p(ids, C.c_int32), self.n, C.byref(cnt))
        return ids[:cnt.value].copy()

    def apply_post_learning(self) -> None:
        L.call("nbx_net_apply_post_learning", self.h)

    def inject_spikes(self, ids) -> None:
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        L.call("nbx_net_inject_spikes", self.h, _p(ids, C.c_int32), int(ids.shape[0]))

    def commit_firings(self) -> None:
        L.call("nbx_net_commit_firings", self.h)

    def set_output_mask(self, mask) -> None:
        """ConfigureOutputNeuronMask: None marks every neuron as an output."""
        if mask is None:
            L.call("nbx_net_set_output_mask", self.h, None)
            return
        self._mask = np.ascontiguousarray(mask, dtype=np.uint8)
        assert self._mask.shape == (self.n,)
        L.call("nbx_net_set_output_mask", self.h, _p(self._mask, C.c_uint8))

    def neuron_step_outputs(self):
        """Neuron update of the current step; (output firings, all firings), each ascending."""
        out = np.zeros(self.n, dtype=np.int32)
        full = np.zeros(self.n, dtype=np.int32)
        oc, fc = C.c_int32(), C.c_int32()
        L.call("nbx_net_neuron_step_outputs", self.h, _p(out, C.c_int32), self.n, C.byref(oc), _p(full, C.c_int32),
               self.n, C.byref(fc))
        return out[:oc.value].copy(), full[:fc.value].copy()

    def set_interface_targets(self, ids) -> None:
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        L.call("nbx_net_set_interface_targets", self.h, _p(ids, C.c_int32), int(ids.shape[0]))

    def upload_interface_values(self, records) -> None:
        """records: (interface_slot, target_neuron, pending_channel, value, scale, overwrite) tuples."""
        arr = (L.InterfaceValue * max(1, len(records)))()
        for k, (slot, tgt, ch, val, sc, ow) in enumerate(records):
            arr[k] = L.InterfaceValue(int(slot), int(tgt), int(ch), int(bool(ow)), float(val), float(sc))
        L.call("nbx_net_upload_interface_values", self.h, arr, len(records))

    def apply_interface(self) -> None:
        L.call("nbx_net_apply_interface", self.h)

    def upload_interface_values_device(self, d_values: int, specs) -> None:
        """specs: (interface_slot, target_neuron, pending_channel, source_index, scale, overwrite)
        tuples; d_values is a device pointer to the source floats."""
        arr = (L.InterfaceCopy * max(1, len(specs)))()
        for k, (slot, tgt, ch, src, sc, ow) in enumerate(specs):
            arr[k] = L.InterfaceCopy(int(slot), int(tgt), int(ch), int(bool(ow)), int(src), float(sc))
        ptr = d_values if isinstance(d_values, C.c_void_p) else C.c_void_p(int(d_values))
        L.call("nbx_net_upload_interface_values_device", self.h, ptr, arr, len(specs))

    def add_arrivals(self, targets, types, weights, t: int) -> None:
        tg = np.ascontiguousarray(targets, dtype=np.int32)
        ty = np.ascontiguousarray(types, dtype=np.uint8)
        w = np.ascontiguousarray(weights, dtype=np.float32)
        assert tg.shape == ty.shape == w.shape
        L.call("nbx_net_add_arrivals", self.h, _p(tg, C.c_int32), _p(ty, C.c_uint8), _p(w, C.c_float),
               int(tg.shape[0]), int(t))

    def set_current_synapses(self, src, dst, delay) -> None:
        """Type-3 current synapses src[k] -> dst[k] after delay[k] steps (slots in this order)."""
        self._cs = [np.ascontiguousarray(src, dtype=np.int32), np.ascontiguousarray(dst, dtype=np.int32),
                    np.ascontiguousarray(delay, dtype=np.uint8)]
        a, b, d = self._cs
        assert a.size == b.size == d.size
        L.call("nbx_net_set_current_synapses", self.h, a.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p),
               d.ctypes.data_as(C.c_void_p), int(a.size))

    def inject_currents(self, src, current, t: int) -> None:
        """Source neurons src[k] emit current[k] at step t (InputCurrent)."""
        a = np.ascontiguousarray(src, dtype=np.int32)
        c = np.ascontiguousarray(current, dtype=np.float32)
        assert a.size == c.size
        L.call("nbx_net_inject_currents", self.h, a.ctypes.data_as(C.c_void_p), c.ctypes.data_as(C.c_void_p),
               int(a.size), int(t))

    def reset_state(self) -> None:
        L.call("nbx_net_reset_state", self.h)

    def download_pending(self) -> np.ndarray:
        slots, ch = C.c_int32(), C.c_int32()
        L.call("nbx_net_ring_shape", self.h, C.byref(slots), C.byref(ch))
        out = np.zeros((ch.value, self.n), dtype=np.float32)
        cc, cs = C.c_int32(), C.c_int32()
        L.call("nbx_net_download_pending", self.h, _p(out, C.c_float), out.size, C.byref(cc), C.byref(cs))
        return out

    def history(self, t: int) -> np.ndarray:
        ids = np.zeros(self.n, dtype=np.int32)
        cnt = C.c_int32()
        L.call("nbx_net_get_history", self.h, int(t), _p(ids, C.c_int32), self.n, C.byref(cnt))
        return ids[:cnt.value].copy()

    # ------------------------------------------------------------ buffers
    def ring(self) -> np.ndarray:
        slots, ch = C.c_int32(), C.c_int32()
        L.call("nbx_net_ring_shape", self.h, C.byref(slots), C.byref(ch))
        out = np.zeros((slots.value, ch.value, self.n), dtype=np.int32)
        L.call("nbx_net_get_ring", self.h, _p(out, C.c_int32), out.size)
        return out

    def weights(self) -> np.ndarray:
        out = np.zeros(self.syn_count, dtype=np.float32)
        L.call("nbx_net_get_weights", self.h, _p(out, C.c_float), out.size)
        return out

    def set_weights(self, w: np.ndarray) -> None:
        w = np.ascontiguousarray(w, dtype=np.float32)
        L.call("nbx_net_set_weights", self.h, _p(w, C.c_float), w.size)

    def synapse_weight(self, j: int) -> float:
        out = C.c_float()
        L.call("nbx_net_get_synapse_weight", self.h, int(j), C.byref(out))
        return float(out.value)

    def set_synapse_weight(self, j: int, w: float) -> None:
        L.call("nbx_net_set_synapse_weight", self.h, int(j), float(w))

    def stdp_state(self):
        S = self.syn_count
        apre = np.zeros(S, np.float32)
        apost = np.zeros(S, np.float32)
        last = np.zeros(S, np.int32)
        L.call("nbx_net_get_stdp_state", self.h, _p(apre, C.c_float), _p(apost, C.c_float), _p(last, C.c_int32), S)
        return apre, apost, last

    def replay(self, spikes_per_step: list, t0: int = 0) -> np.ndarray:
        """Deliver injected spikes of steps t0..t0+K-1; returns arrivals [K][ch][n] of steps t0+1..t0+K."""
        K = len(spikes_per_step)
        tp = np.zeros(K + 1, dtype=np.int32)
        tp[1:] = np.cumsum([len(s) for s in spikes_per_step])
        ids = np.concatenate([np.asarray(s, dtype=np.int32) for s in spikes_per_step]) if tp[-1] else np.zeros(1, np.int32)
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        out = np.zeros((K, self.spec.channels, self.n), dtype=np.int32)
        L.call("nbx_replay_deliver", self.h, _p(ids, C.c_int32), _p(tp, C.c_int32), int(t0), K, _p(out, C.c_int32))
        return out

    def replay_stdp(self, spikes_per_step: list, t0: int = 0, lead: int = 0) -> int:
        """Advance the plastic synapses through the steps t0 .. t0+K-1-lead of a recorded raster.

        spikes_per_step[k] lists the neurons that fired at step t0 - lead + k; the first `lead`
        lists only supply arrivals (their post-side updates were applied before). Weights and trace
        records change as closed-loop steps with exactly these spikes would change them. Returns
        the number of plastic updates applied."""
        K = len(spikes_per_step)
        tp = np.zeros(K + 1, dtype=np.int32)
        tp[1:] = np.cumsum([len(s) for s in spikes_per_step])
        ids = np.concatenate([np.asarray(s, dtype=np.int32) for s in spikes_per_step]) if tp[-1] else np.zeros(1, np.int32)
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        upd = C.c_int64()
        L.call("nbx_replay_stdp", self.h, _p(ids, C.c_int32), _p(tp, C.c_int32), int(t0), int(lead), K - int(lead),
               C.byref(upd))
        return upd.value

    @staticmethod
    def _monitor_desc(output_dir, sample_interval, flush_interval, spikes, state, weights, pending, neuron_ids, name):
        ids = None if neuron_ids is None else np.ascontiguousarray(neuron_ids, dtype=np.int32)
        d = L.MonitorDesc(output_dir=os.fsencode(output_dir), component_name=name.encode(),
                          sample_interval_steps=sample_interval, flush_interval_steps=flush_interval,
                          record_spikes=int(spikes), record_state=int(state), record_weights=int(weights),
                          record_pending_channels=int(pending),
                          neuron_ids=_p(ids, C.c_int32) if ids is not None else None,
                          neuron_count=0 if ids is None else int(ids.size))
        return d, ids

    def export_monitor(self, output_dir: str, *, sample_interval: int = 1, flush_interval: int = 100,
                       spikes: bool = True, state: bool = True, weights: bool = False, pending: bool = True,
                       neuron_ids=None, name: str = "main") -> None:
        """Writes the DebugMonitor directory (meta.json + CSV tables) the reference's results.py reads:
        spikes from the recorded raster, state / pending / weights as one snapshot."""
        d, _keep = self._monitor_desc(output_dir, sample_interval, flush_interval, spikes, state, weights, pending,
                                      neuron_ids, name)
        L.call("nbx_net_export_monitor", self.h, C.byref(d))

    def monitor_start(self, output_dir: str, *, sample_interval: int = 1, flush_interval: int = 100,
                      spikes: bool = True, state: bool = True, weights: bool = False, pending: bool = True,
                      neuron_ids=None, name: str = "main") -> None:
        """Starts the time-series monitor (SimulationDebugMonitor): every later step t with
        t % sample_interval == 0 is sampled after its update and written to the DebugMonitor tables."""
        d, _keep = self._monitor_desc(output_dir, sample_interval, flush_interval, spikes, state, weights, pending,
                                      neuron_ids, name)
        L.call("nbx_net_monitor_start", self.h, C.byref(d))

    def monitor_flush(self) -> None:
        L.call("nbx_net_monitor_flush", self.h)

    def monitor_stop(self) -> None:
        L.call("nbx_net_monitor_stop", self.h)

    def monitor_frames(self) -> tuple:
        """(frames written, frames held on the device)."""
        w, h = C.c_int64(), C.c_int32()
        L.call("nbx_net_monitor_frames", self.h, C.byref(w), C.byref(h))
        return int(w.value), int(h.value)

    def replay_path(self) -> int:
        """Replay kernel this network uses: REPLAY_GLOBAL, REPLAY_OFFSETS, REPLAY_SLOTS or REPLAY_STREAM."""
        p = C.c_int32()
        L.call("nbx_replay_path", self.h, C.byref(p))
        return p.value

    def close(self) -> None:
        if getattr(self, "h", None):
            L.lib().nbx_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def nbnet_info(path: str, verify_checksum: bool = True) -> "L.NbnetHeader":
    """Validated header of a .nbnet file (and its SHA-256 when verify_checksum)."""
    h = L.NbnetHeader()
    L.call("nbx_nbnet_info", os.fsencode(path), int(bool(verify_checksum)), C.byref(h))
    return h


def nbnet_metadata(path: str) -> bytes:
    """The metadata JSON of a .nbnet file."""
    n = C.c_int64()
    L.call("nbx_nbnet_metadata", os.fsencode(path), None, 0, C.byref(n))
    buf = C.create_string_buffer(max(1, n.value))
    L.call("nbx_nbnet_metadata", os.fsencode(path), buf, n.value, C.byref(n))
    return buf.raw[:n.value]
