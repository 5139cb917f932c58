"""Pins the oracle's trigger-driven learning rules (legacy R_STDP and AdditiveKernalChange) with
closed forms on tiny networks whose spike times are forced by current pulses.

Reference paths: L/ = src/native/legacy/source_file_realtime_v1_async/LearningRule/src/Rule/.
"""
import math

import numpy as np
import pytest

from harness import OracleRun
from neuronbridge_amd import _lib as L
from neuronbridge_amd.network import ModelSpec, Synapses
from neuronbridge_amd.workloads import Workload

DT = 0.5
RSTDP = dict(kind=L.RULE_RSTDP, max_ltp=0.92, ltp_tau=16.8, max_ltd=0.53, ltd_tau=33.1, reward=0.01,
             punishment=-0.02, clear_eligibility=1)
ADD = dict(kind=L.RULE_ADDITIVE, ltp_tau=16.8, a1pre=0.002, a2prepre=-0.08)


def forced(n, fire, edges, rules, T):
    """LIF voltage-jump neurons that fire exactly at the steps listed in fire[i] (a current pulse
    that crosses threshold in one step; no other drive). edges: (pre, post, delay, type, weight,
    rule, trigger)."""
    spec = ModelSpec(model=L.LIF_VJ, n=n, dt=DT, t_ref=0, v_rest=-65.0, tau=20.0, v_th=-50.0, r=1.0,
                     v_reset=-65.0)
    edges = sorted(edges, key=lambda e: (e[0], e[2], e[1]))
    row = np.zeros(n + 1, np.int64)
    for e in edges:
        row[e[0] + 1] += 1
    row = np.cumsum(row)
    syn = Synapses(row_ptr=row, post=np.array([e[1] for e in edges], np.int32),
                   delay=np.array([e[2] for e in edges], np.uint8), type=np.array([e[3] for e in edges], np.uint8),
                   weight=np.array([e[4] for e in edges], np.float32), max_delay=max(e[2] for e in edges))
    cur = np.zeros((T, n), np.float32)
    for i, ts in fire.items():
        for t in ts:
            cur[t, i] = 1.0e4
    x0 = np.zeros((n, 2), np.float32)
    x0[:, 0] = -65.0
    wl = Workload("forced", spec, syn=syn, cur=cur, x0=x0, last0=np.full(n, 10000, np.int32))
    wl.rules = {"rules": rules, "rule": np.array([e[5] for e in edges], np.int32),
                "trigger": np.array([e[6] for e in edges], np.int32),
                "max_weight": np.full(len(edges), 1.0, np.float32)}
    return wl, edges


def run(wl, T):
    r = OracleRun(wl)
    _, ras = r.run(T)
    bits = np.unpackbits(ras.view(np.uint8), axis=1, bitorder="little")[:, :wl.spec.n]
    return r, bits


def decay(steps, tau):
    return math.exp(-steps * DT / tau)


def test_forced_spike_times():
    wl, _ = forced(3, {0: [3, 9], 1: [5], 2: [12]}, [(0, 1, 2, 0, 0.0, -1, -1)], [RSTDP], 20)
    _, bits = run(wl, 20)
    assert [list(np.nonzero(bits[:, i])[0]) for i in range(3)] == [[3, 9], [5], [12]]


def test_rstdp_pre_post_trigger_closed_form():
    # pre 0 fires at 4 (arrives at 6 over delay 2), post 1 fires at 10, reward neuron 2 fires at 20
    # (arrives at 21): e = Max_LTP * exp(-(10-6) dt / tau+), then w += reward * e and e is cleared
    w0 = 0.2
    edges = [(0, 1, 2, 0, w0, 0, -1), (2, 1, 1, 0, 0.0, -1, 0)]
    wl, edges = forced(3, {0: [4], 1: [10], 2: [20]}, edges, [RSTDP], 30)
    r, _ = run(wl, 30)
    e = RSTDP["max_ltp"] * decay(10 - 6, RSTDP["ltp_tau"])
    j = next(k for k, x in enumerate(edges) if x[5] == 0)
    assert abs(r.w[j] - (w0 + RSTDP["reward"] * e)) < 1e-6
    assert r.elig[j] == 0.0
    # traces were brought to the trigger step: Apre and Apost both decayed to step 21
    assert abs(r.apre[j] - decay(21 - 6, RSTDP["ltp_tau"])) < 1e-6
    assert abs(r.apost[j] - decay(21 - 10, RSTDP["ltd_tau"])) < 1e-6


def test_rstdp_post_then_pre_depresses_eligibility_and_punishment_uses_inhibitory_trigger():
    # post at 3, pre arrival at 8: e = -Max_LTD * exp(-5 dt / tau-); an inhibitory trigger arriving
    # at 15 applies the punishment factor
    w0 = 0.5
    edges = [(0, 1, 1, 0, w0, 0, -1), (2, 1, 1, 1, 0.0, -1, 0)]
    wl, edges = forced(3, {0: [7], 1: [3], 2: [14]}, edges, [RSTDP], 20)
    r, _ = run(wl, 20)
    e = -RSTDP["max_ltd"] * decay(8 - 3, RSTDP["ltd_tau"])
    j = next(k for k, x in enumerate(edges) if x[5] == 0)
    assert abs(r.w[j] - (w0 + RSTDP["punishment"] * e)) < 1e-6


@pytest.mark.parametrize("reward_first", [True, False])
def test_same_step_triggers_apply_in_ascending_synapse_order(reward_first):
    # two triggers into one target arrive in the same step; with the eligibility cleared by the
    # first, only the lower-numbered trigger synapse changes the weight
    w0 = 0.4
    rw, pn = (2, 3) if reward_first else (3, 2)
    edges = [(0, 1, 1, 0, w0, 0, -1), (rw, 1, 1, 0, 0.0, -1, 0), (pn, 1, 1, 1, 0.0, -1, 0)]
    wl, edges = forced(4, {0: [2], 1: [6], 2: [11], 3: [11]}, edges, [RSTDP], 16)
    r, _ = run(wl, 16)
    e = RSTDP["max_ltp"] * decay(6 - 3, RSTDP["ltp_tau"])
    j = next(k for k, x in enumerate(edges) if x[5] == 0)
    factor = RSTDP["reward"] if reward_first else RSTDP["punishment"]
    assert abs(r.w[j] - (w0 + factor * e)) < 1e-6


def test_rstdp_without_clearing_accumulates_and_clamps():
    rule = dict(RSTDP, clear_eligibility=0, reward=2.0)
    edges = [(0, 1, 1, 0, 0.1, 0, -1), (2, 1, 1, 0, 0.0, -1, 0)]
    wl, edges = forced(3, {0: [2], 1: [5], 2: [8, 10]}, edges, [rule], 14)
    r, _ = run(wl, 14)
    j = next(k for k, x in enumerate(edges) if x[5] == 0)
    e = RSTDP["max_ltp"] * decay(5 - 3, RSTDP["ltp_tau"])
    assert r.elig[j] == np.float32(e) or abs(r.elig[j] - e) < 1e-6
    assert r.w[j] == np.float32(1.0)  # 0.1 + 2 e twice, bounded by max_weight


def test_additive_kernel_closed_form():
    # pre arrivals at 3 and 7 each add a1pre; the trigger arriving at 12 adds
    # a2prepre * (exp(-(12-3) dt / tau) + exp(-(12-7) dt / tau))
    w0 = 0.3
    edges = [(0, 1, 1, 0, w0, 0, -1), (2, 1, 1, 0, 0.0, -1, 0)]
    wl, edges = forced(3, {0: [2, 6], 2: [11]}, edges, [ADD], 16)
    r, _ = run(wl, 16)
    j = next(k for k, x in enumerate(edges) if x[5] == 0)
    apre = decay(12 - 3, ADD["ltp_tau"]) + decay(12 - 7, ADD["ltp_tau"])
    assert abs(r.w[j] - (w0 + 2 * ADD["a1pre"] + ADD["a2prepre"] * apre)) < 1e-6
    assert r.apost[j] == 0.0 and r.elig[j] == 0.0  # the additive rule keeps only Apre
    # the trigger synapse and non-plastic synapses keep their weights
    assert all(r.w[k] == np.float32(x[4]) for k, x in enumerate(edges) if x[5] < 0)


def test_additive_kernel_floor_at_zero():
    rule = dict(ADD, a2prepre=-50.0)
    edges = [(0, 1, 1, 0, 0.05, 0, -1), (2, 1, 1, 0, 0.0, -1, 0)]
    wl, edges = forced(3, {0: [2], 2: [4]}, edges, [rule], 8)
    r, _ = run(wl, 8)
    j = next(k for k, x in enumerate(edges) if x[5] == 0)
    assert r.w[j] == 0.0


def test_trigger_reaches_only_its_rule_and_target():
    # target 1 has one input of each rule; a rule-1 trigger changes only the rule-1 input; the
    # rule-0 input of another target (3) is untouched by triggers into 1
    edges = [(0, 1, 1, 0, 0.3, 0, -1), (4, 1, 1, 0, 0.3, 1, -1), (0, 3, 1, 0, 0.3, 0, -1),
             (2, 1, 1, 0, 0.0, -1, 1)]
    wl, edges = forced(5, {0: [2], 4: [2], 1: [5], 3: [5], 2: [9]}, edges, [RSTDP, ADD], 14)
    r, _ = run(wl, 14)
    by = {(x[0], x[1], x[5]): k for k, x in enumerate(edges)}
    assert r.elig[by[(0, 1, 0)]] != 0.0 and r.w[by[(0, 1, 0)]] == np.float32(0.3)
    assert r.elig[by[(0, 3, 0)]] != 0.0 and r.w[by[(0, 3, 0)]] == np.float32(0.3)
    want = 0.3 + ADD["a1pre"] + ADD["a2prepre"] * decay(10 - 3, ADD["ltp_tau"])
    assert abs(r.w[by[(4, 1, 1)]] - want) < 1e-6


# ------------------------------------------------------------------ cerebellar (L/CerebullarLearningRule.cpp)

import ctypes as C  # noqa: E402

import oracle as O  # noqa: E402

CB = dict(kind=L.RULE_CEREBELLAR, a1pre=0.01, a2prepre=-0.05, initpos=5.0, maxpos=10.0, kernel_step=1.0,
          min_weight=0.05, seed=17)


def orc_rule(d):
    r = O.OrcRule()
    r.kind = d["kind"]
    for f in ("a1pre", "a2prepre", "initpos", "maxpos", "kernel_step", "min_weight"):
        setattr(r, f, d.get(f, 0.0))
    r.seed = d.get("seed", 0)
    return r


def kernel(d):
    r = orc_rule(d)
    mt, inv = C.c_float(), C.c_float()
    n = O.lib().orc_cerebellar_kernel(C.byref(r), None, 0, C.byref(mt), C.byref(inv))
    tab = np.zeros(n, np.float32)
    O.lib().orc_cerebellar_kernel(C.byref(r), O.ptr(tab, C.c_float), n, None, None)
    return tab, mt.value, inv.value


def test_cerebellar_kernel_shape_and_window():
    # alpha function of (t - initpos) / width peaking at 1 exactly at maxpos, zero before initpos;
    # the window ends where x e^(1-x) first falls below 0.01 (x in widths past maxpos' start)
    d = dict(CB, initpos=120.0, maxpos=150.0, kernel_step=0.1)
    tab, mt, inv = kernel(d)
    assert abs(inv - 10.0) < 1e-5
    assert np.all(tab[:1200] == 0.0)
    assert abs(tab[1500] - 1.0) < 1e-5 and tab.argmax() in (1499, 1500, 1501)
    from scipy.optimize import brentq
    x = brentq(lambda v: v * math.exp(1.0 - v) - 0.01, 1.0, 20.0)
    assert abs(mt - (120.0 + 30.0 * x)) < 0.11
    assert len(tab) == int(mt * 10.0) + 1


def test_stable_unit_random():
    f = O.lib().orc_stable_unit_random
    u = np.array([f(17, 2, s, t) for s in range(200) for t in range(100)], np.float64)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert np.all(np.abs(u * 2 ** 24 - np.rint(u * 2 ** 24)) == 0)  # 24-bit values
    hist = np.histogram(u, bins=20, range=(0, 1))[0]
    assert ((hist - 1000.0) ** 2 / 1000.0).sum() < 50  # chi-square, 19 dof
    base = f(17, 2, 5, 9)
    assert len({base, f(18, 2, 5, 9), f(17, 3, 5, 9), f(17, 2, 6, 9), f(17, 2, 5, 10)}) == 5


def test_cerebellar_pre_and_trigger_closed_form():
    # dt 1 ms; arrivals at 3 and 8 each add a1pre * U(seed, rule, connection, step); a trigger
    # arriving at 20 adds a2prepre * kernel(20 - t) for both, newest first, bounded below
    global DT
    old = DT
    DT = 1.0
    try:
        w0 = 0.5
        edges = [(0, 1, 1, 0, w0, 0, -1), (2, 1, 1, 0, 0.0, -1, 0)]
        wl, edges = forced(3, {0: [2, 7], 2: [19]}, edges, [CB], 30)
        r, _ = run(wl, 30)
    finally:
        DT = old
    j = next(k for k, x in enumerate(edges) if x[5] == 0)
    f = O.lib().orc_stable_unit_random
    u = f(17, 0, j, 3) + f(17, 0, j, 8)
    k = lambda e: (1 / 5.0) * (e - 5.0) * math.exp(-(e - 5.0) / 5.0 + 1.0)
    want = w0 + CB["a1pre"] * u + CB["a2prepre"] * (k(12.0) + k(17.0))
    assert abs(r.w[j] - want) < 1e-5


def test_cerebellar_window_and_floor():
    global DT
    old = DT
    DT = 1.0
    try:
        # arrivals far older than the window: the trigger leaves the weight alone
        edges = [(0, 1, 1, 0, 0.5, 0, -1), (2, 1, 1, 0, 0.0, -1, 0)]
        wl, edges = forced(3, {0: [2], 2: [90]}, edges, [CB], 95)
        r, _ = run(wl, 95)
        j = next(k for k, x in enumerate(edges) if x[5] == 0)
        assert abs(r.w[j] - (0.5 + CB["a1pre"] * O.lib().orc_stable_unit_random(17, 0, j, 3))) < 1e-6
        # a large depression stops at min_weight
        wl, edges = forced(3, {0: [2], 2: [12]}, edges, [dict(CB, a2prepre=-100.0)], 16)
        r, _ = run(wl, 16)
        assert r.w[j] == np.float32(CB["min_weight"])
    finally:
        DT = old
