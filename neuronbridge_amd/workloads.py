"""Synthetic, seeded workloads of the benchmark configurations (inputs only).

Each function returns a ``Workload``: model parameters, synapses, input settings
and the initial state, at the configuration's full size by default and at any
smaller size for tests. Seeds are ``0x5EED0000 + config number``.

  sparse_lif  (config 2) LIF voltage-jump, fixed in-degree CSR, integer delays
              U{1..20}, w_E = +0.1 mV, w_I = -0.5 mV (stored as type 1, +0.5),
              tau_m 20 ms, V_th -50, V_reset = V_rest -65, t_ref 2 ms, dt 0.1 ms,
              Poisson background 1000 sources x 10 Hz -> lambda = 1.0 per step, w 0.1 mV
  dense_izh   (config 1) Izhikevich 800 RS / 200 FS, dense all-to-all signed
              weights (exc U[0,0.5], inh U[-1,0]), current input 5N(0,1) / 2N(0,1)
              per step, dt 0.1 ms
  stdp_izh    (config 3) current-based Izhikevich RS/FS 80/20 (the polychronization
              network form: each arrival adds its signed weight to one step's current),
              in-degree 500, E->E plastic pair STDP (A+ 0.01, A- 0.012, tau 20 ms,
              w in [0,1], initial U[0,0.5]), inhibitory U[-1,0], delays U{1..10},
              Poisson input at 5 Hz from 1000 sources per neuron (w 1.3), dt 0.5 ms
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .network import ModelSpec, Net, Synapses, build_fixed_indegree

SEED_BASE = 0x5EED0000


@dataclass
class Workload:
    name: str
    spec: ModelSpec
    syn: Synapses | None = None
    dense_wt: np.ndarray | None = None
    dense_delay: int = 1
    pois_seed: int = 0
    pois_lambda: float = 0.0
    w_bg: float = 0.0
    cur: np.ndarray | None = None
    stdp: dict | None = None
    rules: dict | None = None  # {"rules": [...], "rule", "trigger", "max_weight": per-synapse arrays}
    x0: np.ndarray | None = None
    last0: np.ndarray | None = None
    steps: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def d_slots(self) -> int:
        md = 1
        if self.syn is not None:
            md = max(md, int(self.syn.max_delay))
        if self.dense_wt is not None:
            md = max(md, self.dense_delay)
        return max(md + 1, 3)

    def make_net(self, device: int = 0, fixed_shift: int = 0, nbnet: str | None = None,
                 nbnet_rules: list | None = None) -> Net:
        """The workload as a finalized network; `nbnet` loads the synapses from that .nbnet file
        (the connection table of this workload) instead of the in-memory CSR, and `nbnet_rules`
        then assigns the file's rule ids (nbx_net_assign_nbnet_rules)."""
        net = Net(self.spec, device)
        if nbnet is not None:
            net.load_nbnet(nbnet)
        elif self.syn is not None:
            net.set_sparse(self.syn)
        if self.dense_wt is not None:
            net.set_dense(self.dense_wt, self.dense_delay)
        if self.pois_lambda > 0:
            net.set_poisson(self.pois_seed, self.pois_lambda, self.w_bg)
        if self.stdp is not None and nbnet is None:
            net.set_stdp(**{k: v for k, v in self.stdp.items() if k != "max_weight"})
        if self.rules is not None and nbnet is None:
            net.set_rules(self.rules["rules"], self.rules["rule"], self.rules["trigger"], self.rules["max_weight"])
        if nbnet_rules is not None:
            net.assign_nbnet_rules(nbnet_rules)
        elif self.stdp is not None and nbnet is not None:
            net.set_stdp(**{k: v for k, v in self.stdp.items() if k != "max_weight"})
        if self.cur is not None:
            net.set_current(self.cur, 0)
        if fixed_shift:
            net.set_fixed_shift(fixed_shift)
        net.finalize()
        if self.x0 is not None:
            net.set_state(self.x0, self.last0)
        return net


def sparse_lif(n: int = 100_000, k_in: int = 1000, integrator: int = L.RK4, steps: int = 10_000,
               seed: int = SEED_BASE + 2, threads: int = 0) -> Workload:
    n_exc = int(round(0.8 * n))
    spec = ModelSpec(model=L.LIF_VJ, n=n, dt=0.1, integrator=integrator, t_ref=20,
                     v_rest=-65.0, tau=20.0, v_th=-50.0, r=1.0, v_reset=-65.0)
    syn = build_fixed_indegree(n, k_in, 1, 20, seed, n_exc, threads)
    syn.weight = np.where(syn.type == 0, np.float32(0.1), np.float32(0.5)).astype(np.float32)
    rng = np.random.default_rng(seed)
    x0 = np.zeros((n, 2), dtype=np.float32)
    x0[:, 0] = rng.uniform(-65.0, -50.0, n).astype(np.float32)
    last0 = np.full(n, 10000, dtype=np.int32)
    return Workload("sparse_lif", spec, syn=syn, pois_seed=seed, pois_lambda=1000 * 10.0 * 0.1e-3, w_bg=0.1,
                    x0=x0, last0=last0, steps=steps, extra={"n_exc": n_exc, "k_in": k_in})


def model_sparse(model: int, n: int = 2000, k_in: int = 100, integrator: int = L.EULER, seed: int = SEED_BASE + 6,
                 nmda_fraction: float = 0.3, topology: Synapses | None = None) -> Workload:
    """The config-2 topology (fixed in-degree, delays U{1..20}, 80% excitatory) driving one of the
    remaining legacy models (synaptic weights as listed for in-degree 100, scaled by 100 / k_in so
    the total input per neuron does not depend on the in-degree):
      LIF_EXP2  conductance LIF (V_rest = V_reset -60, tau 20, V_th -50, t_ref 2 ms, Eexc 0 / tau 5,
                Einh -80 / tau 10), g increments 0.02 exc / 0.05 inh, background 0.008
      LIF_EXP3  the same with AMPA / GABA / NMDA (tau 10 each); nmda_fraction of the excitatory
                synapses are NMDA (type 2)
      POISSON   rate 30 Hz + 20 Hz per background event + 200 Hz per excitatory / -400 Hz per
                inhibitory arrival
      TRIGGER_RELAY threshold 0.25; background events of 0.1, excitatory arrivals 0.002"""
    n_exc = int(round(0.8 * n))
    if topology is None:
        syn = build_fixed_indegree(n, k_in, 1, 20, seed, n_exc)
    else:  # the same (n, k_in, seed) topology built once by the caller; types are copied
        syn = Synapses(topology.row_ptr, topology.post, topology.delay, topology.type.copy(), topology.weight,
                       topology.max_delay)
    rng = np.random.default_rng(seed)
    lam = 1.0
    x0 = None
    g = np.float32(100.0 / k_in)
    if model in (L.LIF_EXP2, L.LIF_EXP3):
        spec = ModelSpec(model=model, n=n, dt=0.1, integrator=integrator, t_ref=20, v_rest=-60.0, tau=20.0,
                         v_th=-50.0, r=1.0, v_reset=-60.0, e_ampa=0.0, e_gaba=-80.0,
                         ampa_tau=5.0 if model == L.LIF_EXP2 else 10.0, gaba_tau=10.0, nmda_tau=10.0)
        syn.weight = (np.where(syn.type == 0, np.float32(0.02), np.float32(0.05)) * g).astype(np.float32)
        if model == L.LIF_EXP3:
            nmda = (syn.type == 0) & (rng.random(syn.count) < nmda_fraction)
            syn.type = np.where(nmda, 2, syn.type).astype(np.uint8)
        w_bg = 0.008
        x0 = np.zeros((n, spec.nstate), np.float32)
        x0[:, 0] = rng.uniform(-60.0, -50.0, n).astype(np.float32)
    elif model == L.POISSON:
        spec = ModelSpec(model=model, n=n, dt=0.1, rate_bias_hz=30.0, rate_gain_hz=1.0)
        syn.weight = (np.where(syn.type == 0, np.float32(200.0), np.float32(400.0)) * g).astype(np.float32)
        w_bg = 20.0
    elif model == L.TRIGGER_RELAY:
        spec = ModelSpec(model=model, n=n, dt=0.1, v_th=0.25)
        syn.weight = (np.where(syn.type == 0, np.float32(0.002), np.float32(0.5)) * g).astype(np.float32)
        w_bg = 0.1
    else:
        raise ValueError("model_sparse covers LIF_EXP2, LIF_EXP3, POISSON and TRIGGER_RELAY")
    last0 = np.full(n, 10000, dtype=np.int32) if x0 is not None else None
    return Workload(f"sparse_model_{model}", spec, syn=syn, pois_seed=seed, pois_lambda=lam, w_bg=w_bg, x0=x0,
                    last0=last0, steps=0, extra={"n_exc": n_exc, "k_in": k_in})


def populations(models=(L.LIF_VJ, L.IZH), sizes=(1600, 400), k_in: int = 100, integrator: int = L.EULER,
                seed: int = SEED_BASE + 8, nmda_fraction: float = 0.25) -> Workload:
    """Several populations of different models in one network, joined by one fixed in-degree
    topology (delays U{1..20}; presynaptic neurons of the last 20% of each population inhibitory).
    Each synapse's weight follows its target's model: voltage jumps of 0.12 / 0.45 mV into LIF
    voltage-jump, conductance increments of 0.02 / 0.05 into LIF exp-decay and the double / triple
    exponential LIFs, 0.05 / 0.1 into conductance Izhikevich, signed currents 1.5 / -3 into
    current-based Izhikevich, rate increments of 200 / 400 Hz into PoissonRate, 0.002 / 0.5 into the
    trigger relay (all for in-degree 100, scaled by 100 / k_in). nmda_fraction of the excitatory
    synapses into conductance Izhikevich and the triple-exponential LIF are NMDA (type 2).
    Background: Poisson(1.0) per step of weight 0.1 into every neuron."""
    from .network import PopulationsSpec
    rng = np.random.default_rng(seed)
    n = int(sum(sizes))
    begins = np.cumsum([0] + list(sizes))
    exc = np.ones(n, dtype=bool)
    for b, m in zip(begins[:-1], sizes):
        exc[b + int(round(0.8 * m)):b + m] = False
    syn = build_fixed_indegree(n, k_in, 1, 20, seed, n)  # types set below from each presynaptic neuron
    pre = np.repeat(np.arange(n, dtype=np.int64), np.diff(syn.row_ptr))
    syn.type = np.where(exc[pre], 0, 1).astype(np.uint8)
    g = np.float32(100.0 / k_in)
    pops, x0cols = [], []
    w = np.zeros(syn.count, np.float32)
    tgt_pop = np.searchsorted(begins, syn.post, side="right") - 1
    for k, (model, m) in enumerate(zip(models, sizes)):
        into = tgt_pop == k
        if model == L.LIF_VJ:
            ps = ModelSpec(model=model, n=m, dt=0.1, integrator=integrator, t_ref=20, v_rest=-65.0, tau=20.0,
                           v_th=-50.0, r=1.0, v_reset=-65.0)
            we, wi = 0.12, 0.45
        elif model == L.LIF_EXP:
            ps = ModelSpec(model=model, n=m, dt=0.1, integrator=integrator, t_ref=20, v_rest=-65.0, tau=20.0,
                           v_th=-50.0, r=1.0, v_reset=-65.0, g_tau=5.0, e_rev=0.0)
            we, wi = 0.02, 0.05
        elif model in (L.LIF_EXP2, L.LIF_EXP3):
            ps = ModelSpec(model=model, n=m, dt=0.1, integrator=integrator, t_ref=20, v_rest=-60.0, tau=20.0,
                           v_th=-50.0, r=1.0, v_reset=-60.0, e_ampa=0.0, e_gaba=-80.0, ampa_tau=5.0,
                           gaba_tau=10.0, nmda_tau=10.0)
            we, wi = 0.02, 0.05
        elif model in (L.IZH, L.IZH_CUR):
            a, b, c, d = izh_params(m, int(round(0.8 * m)), rng)
            ps = ModelSpec(model=model, n=m, dt=0.1, integrator=integrator, v_rest=-65.0, a=a, b=b, c=c, d=d,
                           nmda_tau=100.0)
            we, wi = (0.05, 0.1) if model == L.IZH else (1.5, -3.0)
        elif model == L.POISSON:
            ps = ModelSpec(model=model, n=m, dt=0.1, rate_bias_hz=30.0, rate_gain_hz=1.0)
            we, wi = 200.0, 400.0
        else:
            ps = ModelSpec(model=model, n=m, dt=0.1, v_th=0.25)
            we, wi = 0.002, 0.5
        w[into] = np.where(syn.type[into] == 0, np.float32(we), np.float32(wi)) * g
        if model in (L.IZH, L.LIF_EXP3):
            nmda = into & (syn.type == 0) & (rng.random(syn.count) < nmda_fraction)
            syn.type = np.where(nmda, 2, syn.type).astype(np.uint8)
        pops.append(ps)
    syn.weight = w.astype(np.float32)
    spec = PopulationsSpec(pops)
    x0 = np.zeros((n, spec.nstate), np.float32)
    for b, ps in zip(begins[:-1], pops):
        if ps.model in (L.LIF_VJ, L.LIF_EXP):
            x0[b:b + ps.n, 0] = rng.uniform(-65.0, -50.0, ps.n)
        elif ps.model in (L.LIF_EXP2, L.LIF_EXP3):
            x0[b:b + ps.n, 0] = rng.uniform(-60.0, -50.0, ps.n)
        elif ps.model in (L.IZH, L.IZH_CUR):
            x0[b:b + ps.n, 0] = rng.uniform(-70.0, -55.0, ps.n)
            x0[b:b + ps.n, 1] = np.asarray(ps.b, np.float32) * x0[b:b + ps.n, 0]
    last0 = np.full(n, 10000, dtype=np.int32)
    return Workload("populations", spec, syn=syn, pois_seed=seed, pois_lambda=1.0, w_bg=0.1, x0=x0, last0=last0,
                    steps=0, extra={"begins": begins.tolist(), "k_in": k_in})


def trigger_rules(n: int = 2000, k_in: int = 100, seed: int = SEED_BASE + 7, trigger_fraction: float = 0.03,
                  with_stdp: bool = False, cerebellar: bool = False, history: int = 0) -> Workload:
    """The config-2 LIF network with the legacy trigger-driven rules on its E->E synapses:
    40% R-STDP (rule 0: Max_LTP 0.92, LTP_tau 16.8, Max_LTD 0.53, LTD_tau 33.1, reward 0.01,
    punishment -0.01, eligibility cleared after a trigger), 40% additive kernel (rule 1: a1pre 5e-4,
    a2prepre -0.02, LTP_tau 16.8), weights bounded by 0.3; trigger_fraction of all synapses are
    triggers (half of each rule; inhibitory trigger synapses punish). with_stdp makes the remaining
    E->E synapses pair-STDP plastic. cerebellar moves a third of the rule synapses and triggers to a
    cerebellar rule (rule 2: alpha 2e-4, beta -5e-3, kernel from 2 to 6 ms in 0.1 ms steps, floor
    0.01, seed 17, `history` arrivals kept per synapse)."""
    wl = sparse_lif(n=n, k_in=k_in, steps=0, seed=seed)
    s = wl.syn
    rng = np.random.default_rng(seed)
    n_exc = wl.extra["n_exc"]
    pre = np.repeat(np.arange(n, dtype=np.int64), np.diff(s.row_ptr))
    ee = (pre < n_exc) & (s.post < n_exc)
    u = rng.random(s.count)
    rule = np.full(s.count, -1, np.int32)
    nr = 3 if cerebellar else 2
    for r in range(nr):
        rule[ee & (u >= 0.8 * r / nr) & (u < 0.8 * (r + 1) / nr)] = r
    trig = np.full(s.count, -1, np.int32)
    t_sel = (rule < 0) & (rng.random(s.count) < trigger_fraction)
    trig[t_sel] = rng.integers(0, nr, int(t_sel.sum()))
    if with_stdp:
        s.plastic = ((rule < 0) & (trig < 0) & ee).astype(np.uint8)
        wl.stdp = dict(max_ltp=0.001, ltp_tau=20.0, max_ltd=0.0012, ltd_tau=20.0, w_max=0.3)
    rules = [dict(kind=L.RULE_RSTDP, max_ltp=0.92, ltp_tau=16.8, max_ltd=0.53, ltd_tau=33.1, reward=0.01,
                  punishment=-0.01, clear_eligibility=1),
             dict(kind=L.RULE_ADDITIVE, ltp_tau=16.8, a1pre=5e-4, a2prepre=-0.02)]
    if cerebellar:
        rules.append(dict(kind=L.RULE_CEREBELLAR, a1pre=2e-4, a2prepre=-5e-3, initpos=2.0, maxpos=6.0,
                          kernel_step=0.1, min_weight=0.01, seed=17, history=history))
    wl.rules = {"rules": rules, "rule": rule, "trigger": trig, "max_weight": np.full(s.count, 0.3, np.float32)}
    wl.name = "trigger_rules"
    return wl


def izh_params(n: int, n_exc: int, rng) -> tuple:
    """Izhikevich (2003) RS/FS parameter draws: RS a=.02 b=.2 c=-65+15r^2 d=8-6r^2; FS a=.02+.08r b=.25-.05r c=-65 d=2."""
    re = rng.random(n_exc)
    ri = rng.random(n - n_exc)
    a = np.concatenate([np.full(n_exc, 0.02), 0.02 + 0.08 * ri]).astype(np.float32)
    b = np.concatenate([np.full(n_exc, 0.2), 0.25 - 0.05 * ri]).astype(np.float32)
    c = np.concatenate([-65.0 + 15.0 * re ** 2, np.full(n - n_exc, -65.0)]).astype(np.float32)
    d = np.concatenate([8.0 - 6.0 * re ** 2, np.full(n - n_exc, 2.0)]).astype(np.float32)
    return a, b, c, d


def dense_izh(n: int = 1000, steps: int = 10_000, integrator: int = L.RK4, seed: int = SEED_BASE + 1,
              noise_scale: float = 1.0) -> Workload:
    """noise_scale multiplies the per-step current draws (1.0 = the configuration as stated; at dt 0.1 ms
    sqrt(10) gives the per-millisecond noise variance of the original 1 ms Izhikevich network)."""
    n_exc = int(round(0.8 * n))
    rng = np.random.default_rng(seed)
    a, b, c, d = izh_params(n, n_exc, rng)
    spec = ModelSpec(model=L.IZH_CUR, n=n, dt=0.1, integrator=integrator, v_rest=-65.0, izh_v_th=30.0, izh_r=1.0,
                     a=a, b=b, c=c, d=d)
    wt = np.empty((n, n), dtype=np.float32)  # wt[pre][post]
    wt[:n_exc] = rng.uniform(0.0, 0.5, (n_exc, n)).astype(np.float32)
    wt[n_exc:] = rng.uniform(-1.0, 0.0, (n - n_exc, n)).astype(np.float32)
    cur = np.empty((steps, n), dtype=np.float32)
    cur[:, :n_exc] = (noise_scale * 5.0 * rng.standard_normal((steps, n_exc))).astype(np.float32)
    cur[:, n_exc:] = (noise_scale * 2.0 * rng.standard_normal((steps, n - n_exc))).astype(np.float32)
    return Workload("dense_izh", spec, dense_wt=wt, dense_delay=1, cur=cur, steps=steps, extra={"n_exc": n_exc})


def stdp_izh(n: int = 50_000, k_in: int = 500, steps: int = 20_000, integrator: int = L.EULER,
             seed: int = SEED_BASE + 3, pois_rate_hz: float = 5.0, pois_sources: int = 1000, w_bg: float = 1.3,
             w_inh: float = 1.0, threads: int = 0) -> Workload:
    n_exc = int(round(0.8 * n))
    rng = np.random.default_rng(seed)
    a, b, c, d = izh_params(n, n_exc, rng)
    dt = 0.5
    spec = ModelSpec(model=L.IZH_CUR, n=n, dt=dt, integrator=integrator, v_rest=-65.0, izh_v_th=30.0, izh_r=1.0,
                     a=a, b=b, c=c, d=d)
    syn = build_fixed_indegree(n, k_in, 1, 10, seed, n_exc, threads)
    S = syn.count
    pre = np.repeat(np.arange(n, dtype=np.int32), np.diff(syn.row_ptr))
    w = rng.uniform(0.0, 0.5, S).astype(np.float32)
    inh = syn.type == 1
    w[inh] = rng.uniform(-w_inh, 0.0, int(inh.sum())).astype(np.float32)  # signed current weights
    syn.weight = w
    syn.plastic = ((pre < n_exc) & (syn.post < n_exc)).astype(np.uint8)
    stdp = dict(max_ltp=0.01, ltp_tau=20.0, max_ltd=0.012, ltd_tau=20.0, w_max=1.0)
    lam = pois_sources * pois_rate_hz * dt * 1e-3
    return Workload("stdp_izh", spec, syn=syn, pois_seed=seed, pois_lambda=lam, w_bg=w_bg, stdp=stdp,
                    steps=steps, extra={"n_exc": n_exc, "k_in": k_in})


# ---------------------------------------------------------------------------------------------
# config 4: batched independent LIF instances with dense BF16 recurrent weights

BATCH_LIF = dict(integrator=L.RK4, dt=0.1, t_ref=20, v_rest=-65.0, v_th=-50.0, r=1.0, v_reset=-65.0, w_bg=0.1)
MAX_ABS_NORMAL = 6.8  # |z| bound of the Box-Muller draws (u1 >= 2^-33)


def batched_lif(count: int = 256, n_lo: int = 512, n_hi: int = 8192, steps: int = 5000,
                seed: int = SEED_BASE + 4, threads: int = 0):
    """256 instances, N ~ U{512..8192}, tau_m ~ U[10,30] ms, input rate ~ U[5,20] Hz (1000 sources, 0.1 mV),
    W ~ N(0, 1/N) rounded to BF16, RK4 dt 0.1 ms. Returns (instances, shared parameters)."""
    from .batch import Instance, dense_normal_bf16, poisson_kmax, shift_for_bound
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(rng.integers(n_lo, n_hi + 1))
        tau = float(rng.uniform(10.0, 30.0))
        rate = float(rng.uniform(5.0, 20.0))
        lam = 1000 * rate * BATCH_LIF["dt"] * 1e-3
        sigma = 1.0 / np.sqrt(n)
        iseed = (seed << 16) + k
        w = dense_normal_bf16(n, sigma, iseed, threads)
        bound = n * MAX_ABS_NORMAL * sigma * 1.01 + poisson_kmax(lam) * BATCH_LIF["w_bg"]
        v0 = rng.uniform(-65.0, -50.0, n).astype(np.float32)
        out.append(Instance(n=n, tau=tau, pois_lambda=lam, seed=iseed, fixed_shift=shift_for_bound(bound),
                            w_bf16=w, v0=v0))
    return out, dict(BATCH_LIF)


def instance_workload(ins, shared) -> Workload:
    """One batched instance as a single-network workload (the plain FP32 per-instance path)."""
    from .batch import bf16_to_f32
    spec = ModelSpec(model=L.LIF_VJ, n=ins.n, dt=shared["dt"], integrator=shared["integrator"], t_ref=shared["t_ref"],
                     v_rest=shared["v_rest"], tau=ins.tau, v_th=shared["v_th"], r=shared["r"],
                     v_reset=shared["v_reset"])
    x0 = np.zeros((ins.n, 2), np.float32)
    x0[:, 0] = ins.v0 if ins.v0 is not None else shared["v_reset"]
    return Workload("batched_instance", spec, dense_wt=bf16_to_f32(ins.w_bf16), dense_delay=1,
                    pois_seed=ins.seed, pois_lambda=ins.pois_lambda, w_bg=shared["w_bg"], x0=x0,
                    last0=np.full(ins.n, 10000, np.int32), extra={"fixed_shift": ins.fixed_shift})


# ---------------------------------------------------------------------------------------------
# config 5: differentiable recurrent LIF trained on sinusoid readout targets

def diff_lif(n: int = 2048, batch: int = 64, steps: int = 500, n_in: int = 200, rate_hz: float = 20.0,
             n_out: int = 8, dt: float = 1.0, seed: int = SEED_BASE + 5):
    """W_rec ~ N(0, 1/N), W_in ~ N(0, 1/n_in), W_out ~ N(0, 1/N); Poisson inputs (rate_hz) as 0/1 per
    (step, channel, trial); targets sin(2 pi f t + phi), f ~ U[0.5, 4] Hz, phi ~ U[0, 2 pi), shared by
    all trials. Returns float32 numpy arrays (W_in, W_rec, W_out, x [T, n_in, B], target [T, n_out, B])."""
    rng = np.random.default_rng(seed)
    W_rec = (rng.standard_normal((n, n)) / np.sqrt(n)).astype(np.float32)
    W_in = (rng.standard_normal((n, n_in)) / np.sqrt(n_in)).astype(np.float32)
    W_out = (rng.standard_normal((n_out, n)) / np.sqrt(n)).astype(np.float32)
    x = (rng.random((steps, n_in, batch)) < rate_hz * dt * 1e-3).astype(np.float32)
    f = rng.uniform(0.5, 4.0, n_out)
    phi = rng.uniform(0.0, 2 * np.pi, n_out)
    t = np.arange(steps) * dt * 1e-3
    tgt = np.sin(2 * np.pi * f[None, :] * t[:, None] + phi[None, :]).astype(np.float32)
    target = np.repeat(tgt[:, :, None], batch, axis=2)
    return W_in, W_rec, W_out, x, target
