"""Test helpers: run a workload through the CPU oracle (the checker)."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

import oracle as O
from neuronbridge_amd import _lib as L
from neuronbridge_amd.network import PopulationsSpec

MODEL_MAP = {L.LIF_VJ: O.LIF_VJ, L.LIF_EXP: O.LIF_EXP, L.IZH: O.IZH, L.IZH_CUR: O.IZH_CUR,
             L.LIF_EXP2: O.LIF_EXP2, L.LIF_EXP3: O.LIF_EXP3, L.POISSON: O.POISSON, L.TRIGGER_RELAY: O.TRIGGER_RELAY}


def auto_shift(wl) -> int:
    """Largest fixed-point shift with no possible int32 overflow (same rule as libnbx)."""
    n = wl.spec.n
    bound = np.zeros(n, dtype=np.float64)
    if wl.syn is not None:
        w = np.abs(wl.syn.weight.astype(np.float64))
        if wl.stdp is not None and wl.syn.plastic is not None:
            wm = wl.stdp.get("max_weight")
            w = np.where(wl.syn.plastic != 0, np.maximum(w, wl.stdp["w_max"] if wm is None else np.abs(wm)), w)
        if wl.rules is not None:
            w = np.where(wl.rules["rule"] >= 0, np.maximum(w, np.abs(wl.rules["max_weight"])), w)
        bound += np.bincount(wl.syn.post, weights=w, minlength=n)
    if wl.dense_wt is not None:
        bound += np.abs(wl.dense_wt.astype(np.float64)).sum(axis=0)
    B = float(bound.max()) if n else 0.0
    if wl.pois_lambda > 0:
        tab = np.zeros(64, dtype=np.uint32)
        kmax = O.lib().orc_poisson_thresholds(wl.pois_lambda, O.ptr(tab, C.c_uint32), 64)
        B += kmax * abs(wl.w_bg)
    s = 24
    while s > 0 and B * math.ldexp(1.0, s) + 1e6 >= 2147483647.0:
        s -= 1
    return s


def initial_state(wl):
    """The state a freshly finalized network holds, then the workload's x0 if any."""
    n, ns = wl.spec.n, wl.spec.nstate
    if wl.x0 is not None:
        return np.ascontiguousarray(wl.x0, dtype=np.float32).reshape(n, ns).copy(), wl.last0.astype(np.int32).copy()
    last = np.full(n, 10000, dtype=np.int32)
    if isinstance(wl.spec, PopulationsSpec):
        x = np.zeros((n, ns), dtype=np.float32)
        for b, ps in zip(wl.spec.begins, wl.spec.pops):
            x[b:b + ps.n, :ps.nstate] = _initial_columns(ps, ps.n)
        return x, last
    return _initial_columns(wl.spec, n), last


def _initial_columns(sp, n):
    x = np.zeros((n, sp.nstate), dtype=np.float32)
    if sp.model in (L.LIF_VJ, L.LIF_EXP3):
        x[:, 0] = sp.v_reset
    elif sp.model in (L.LIF_EXP, L.LIF_EXP2):
        x[:, 0] = sp.v_rest
    elif sp.model in (L.POISSON, L.TRIGGER_RELAY):
        pass
    else:
        x[:, 0] = sp.v_rest
        x[:, 1] = (np.asarray(sp.b, np.float32) * np.float32(sp.v_rest)).astype(np.float32)
    return x


def orc_model(sp, keep: list):
    """An OrcModel of one population's parameters (per-neuron arrays kept alive in `keep`)."""
    m = O.OrcModel(model=MODEL_MAP[sp.model], integrator=sp.integrator, n=sp.n, t_ref=sp.t_ref, dt=sp.dt,
                   v_rest=sp.v_rest, tau=sp.tau, v_th=sp.v_th, r=sp.r, v_reset=sp.v_reset, g_tau=sp.g_tau,
                   e_rev=sp.e_rev, izh_v_th=sp.izh_v_th, izh_r=sp.izh_r, e_ampa=sp.e_ampa, e_gaba=sp.e_gaba,
                   ampa_tau=sp.ampa_tau, gaba_tau=sp.gaba_tau, nmda_tau=sp.nmda_tau,
                   rate_bias=sp.rate_bias_hz, rate_gain=sp.rate_gain_hz)
    if sp.a is not None:
        for name in "abcd":
            arr = np.ascontiguousarray(getattr(sp, name), dtype=np.float32)
            keep.append(arr)
            setattr(m, name, O.ptr(arr, C.c_float))
    return m


def raster_lists(raster: np.ndarray, n: int) -> list:
    """Spike lists per step from a bit raster [steps][ceil(n/32)] (bit i of a word: neuron 32w + i)."""
    bits = np.unpackbits(np.ascontiguousarray(raster).view(np.uint8), axis=1, bitorder="little")[:, :n]
    return [np.nonzero(row)[0].astype(np.int32) for row in bits]


def raster_csr(spikes_per_step: list):
    """(ids, tick_ptr) of spike lists, the layout of the replay entry points."""
    tp = np.zeros(len(spikes_per_step) + 1, dtype=np.int32)
    tp[1:] = np.cumsum([len(s) for s in spikes_per_step])
    ids = (np.concatenate([np.asarray(s, dtype=np.int32) for s in spikes_per_step]) if tp[-1]
           else np.zeros(1, np.int32))
    return np.ascontiguousarray(ids, dtype=np.int32), tp


class OracleRun:
    """Holds an OrcNet built from a workload plus the oracle-side mutable arrays."""

    def __init__(self, wl, fixed_shift: int | None = None):
        self.wl = wl
        sp = wl.spec
        self.keep = []
        k = self.keep.append
        multi = isinstance(sp, PopulationsSpec)
        m = orc_model(sp.pops[0] if multi else sp, self.keep)
        net = O.OrcNet(m=m, d_slots=wl.d_slots, fixed_shift=auto_shift(wl) if fixed_shift is None else fixed_shift)
        if multi:
            pops = (O.OrcModel * len(sp.pops))(*[orc_model(ps, self.keep) for ps in sp.pops])
            begins = np.ascontiguousarray(sp.begins + [sp.n], dtype=np.int32)
            k(pops)
            k(begins)
            net.npops = len(sp.pops)
            net.pops = C.cast(pops, C.POINTER(O.OrcModel))
            net.pop_begin = O.ptr(begins, C.c_int32)
        self.shift = net.fixed_shift
        self.w = None
        if wl.syn is not None:
            s = wl.syn
            self.w = s.weight.astype(np.float32).copy()
            for arr in (s.row_ptr, s.post, s.delay, s.type, self.w):
                k(arr)
            net.row_ptr = O.ptr(s.row_ptr, C.c_int64)
            net.post = O.ptr(s.post, C.c_int32)
            net.delay = O.ptr(s.delay, C.c_uint8)
            net.type = O.ptr(s.type, C.c_uint8)
            net.w = O.ptr(self.w, C.c_float)
        if wl.dense_wt is not None:
            wt = np.ascontiguousarray(wl.dense_wt, dtype=np.float32)
            k(wt)
            net.dense_wt = O.ptr(wt, C.c_float)
            net.dense_delay = wl.dense_delay
        if wl.pois_lambda > 0:
            net.pois_on = 1
            net.seed_lo = wl.pois_seed & 0xFFFFFFFF
            net.seed_hi = (wl.pois_seed >> 32) & 0xFFFFFFFF
            net.pois_lambda = wl.pois_lambda
            net.w_bg = wl.w_bg
        if wl.cur is not None:
            cur = np.ascontiguousarray(wl.cur, dtype=np.float32)
            k(cur)
            net.cur = O.ptr(cur, C.c_float)
            net.cur_t_begin = 0
        if wl.stdp is not None:
            s = wl.syn
            S = s.count
            net.stdp_on = 1
            plastic = np.ascontiguousarray(s.plastic, dtype=np.uint8)
            self.apre = np.zeros(S, np.float32)
            self.apost = np.zeros(S, np.float32)
            self.slast = np.zeros(S, np.int32)
            sel = np.nonzero(plastic)[0].astype(np.int64)
            order = np.argsort(s.post[sel], kind="stable")
            in_syn = np.ascontiguousarray(sel[order])
            in_ptr = np.zeros(sp.n + 1, dtype=np.int64)
            in_ptr[1:] = np.cumsum(np.bincount(s.post[sel], minlength=sp.n))
            for arr in (plastic, self.apre, self.apost, self.slast, in_syn, in_ptr):
                k(arr)
            net.plastic = O.ptr(plastic, C.c_uint8)
            for f in ("max_ltp", "ltp_tau", "max_ltd", "ltd_tau", "w_max"):
                setattr(net, f, wl.stdp[f])
            if wl.stdp.get("max_weight") is not None:  # per-synapse bounds (networks from .nbnet)
                smax = np.ascontiguousarray(wl.stdp["max_weight"], dtype=np.float32)
                k(smax)
                net.stdp_maxw = O.ptr(smax, C.c_float)
            net.apre = O.ptr(self.apre, C.c_float)
            net.apost = O.ptr(self.apost, C.c_float)
            net.slast = O.ptr(self.slast, C.c_int32)
            net.in_ptr = O.ptr(in_ptr, C.c_int64)
            net.in_syn = O.ptr(in_syn, C.c_int64)
        if wl.rules is not None:
            S = wl.syn.count
            rl = wl.rules["rules"]
            arr = (O.OrcRule * len(rl))()
            for kk, r in enumerate(rl):
                arr[kk].kind = r["kind"]
                arr[kk].clear = int(r.get("clear_eligibility", 0))
                for f in ("max_ltp", "ltp_tau", "max_ltd", "ltd_tau", "reward", "punishment", "a1pre", "a2prepre",
                          "initpos", "maxpos", "kernel_step", "min_weight"):
                    setattr(arr[kk], f, float(r.get(f, 0.0)))
                arr[kk].seed = int(r.get("seed", 0))
            k(arr)
            self.syn_rule = np.ascontiguousarray(wl.rules["rule"], dtype=np.int8)
            self.syn_trig = np.ascontiguousarray(wl.rules["trigger"], dtype=np.int8)
            self.maxw = np.ascontiguousarray(wl.rules["max_weight"], dtype=np.float32)
            self.elig = np.zeros(S, np.float32)
            if wl.stdp is None:
                self.apre = np.zeros(S, np.float32)
                self.apost = np.zeros(S, np.float32)
                self.slast = np.zeros(S, np.int32)
                for arr_ in (self.apre, self.apost, self.slast):
                    k(arr_)
                net.apre = O.ptr(self.apre, C.c_float)
                net.apost = O.ptr(self.apost, C.c_float)
                net.slast = O.ptr(self.slast, C.c_int32)
            for arr_ in (self.syn_rule, self.syn_trig, self.maxw, self.elig):
                k(arr_)
            if wl.extra.get("conn_index") is not None:  # connection indices of a .nbnet table
                ci = np.ascontiguousarray(wl.extra["conn_index"], dtype=np.int32)
                k(ci)
                net.conn_index = O.ptr(ci, C.c_int32)
            net.nrules = len(rl)
            net.rules = C.cast(arr, C.c_void_p)
            net.syn_rule = O.ptr(self.syn_rule, C.c_int8)
            net.syn_trig = O.ptr(self.syn_trig, C.c_int8)
            net.maxw = O.ptr(self.maxw, C.c_float)
            net.elig = O.ptr(self.elig, C.c_float)
        self.net = net
        self.x, self.last = initial_state(wl)
        D, n = wl.d_slots, sp.n
        self.hist = np.zeros((D, n), dtype=np.int32)
        self.hist_cnt = np.zeros(D, dtype=np.int32)
        self.ring = np.zeros((D, 3, n), dtype=np.int32)
        self.t = 0
        self.events = np.zeros(1, dtype=np.int64)

    def set_inputs(self, per_step: list, t0: int) -> None:
        """Inputs from outside the synapses for steps t0, t0 + 1, ...: per step a list of
        (target, kind, value) with kind 0..2 an arrival of that synapse type, 3 a new I_ext."""
        ptr = np.zeros(len(per_step) + 1, dtype=np.int32)
        ptr[1:] = np.cumsum([len(x) for x in per_step])
        arr = (O.OrcInput * max(1, int(ptr[-1])))()
        j = 0
        for step in per_step:
            for tgt, kind, val in step:
                arr[j] = O.OrcInput(int(tgt), int(kind), float(np.float32(val)))
                j += 1
        self._inputs = (arr, ptr)
        self.net.inputs = C.cast(arr, C.c_void_p)
        self.net.input_ptr = O.ptr(ptr, C.c_int32)
        self.net.input_t0 = int(t0)
        self.net.input_nsteps = len(per_step)

    def run(self, nsteps: int, acc_mode: int = O.ACC_FIXED, ring: bool = True, raster: bool = True):
        n = self.wl.spec.n
        counts = np.zeros(nsteps, dtype=np.int32)
        words = (n + 31) // 32
        ras = np.zeros((nsteps, words), dtype=np.uint32) if raster else None
        use_ring = ring and acc_mode == O.ACC_FIXED and self.wl.stdp is None and self.wl.rules is None
        st = O.lib().orc_run(C.byref(self.net), acc_mode, O.ptr(self.x, C.c_float), O.ptr(self.last, C.c_int32),
                             O.ptr(self.hist, C.c_int32), O.ptr(self.hist_cnt, C.c_int32),
                             O.ptr(self.ring, C.c_int32) if use_ring else None, self.t, nsteps,
                             O.ptr(counts, C.c_int32), O.ptr(ras, C.c_uint32) if raster else None,
                             O.ptr(self.events, C.c_int64))
        assert st == 0, f"oracle self-check failed (status {st}): push-ring image != pull delivery"
        self.t += nsteps
        return counts, ras

    def replay_stdp(self, spikes_per_step: list, t0: int = 0, lead: int = 0) -> int:
        """orc_replay_stdp on this run's weights and traces; returns the number of updates."""
        ids, tp = raster_csr(spikes_per_step)
        return int(O.lib().orc_replay_stdp(C.byref(self.net), O.ptr(ids, C.c_int32), O.ptr(tp, C.c_int32), int(t0),
                                           int(lead), len(spikes_per_step) - int(lead)))

    def ring_as_device(self, nch: int) -> np.ndarray:
        """Oracle ring image in the device layout [slot][channel][n]."""
        return self.ring[:, :nch, :].copy()
