"""Pins the CPU oracle against the reference's known answers, closed forms and brute force.

Reference paths are relative to the reference root; L/ = src/native/legacy/source_file_realtime_v1_async/.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from harness import OracleRun
from neuronbridge_amd import _lib as L

L_ = O.lib()


def model(**kw):
    base = dict(model=O.LIF_VJ, integrator=O.EULER, n=1, t_ref=50, dt=1.0, v_rest=0.0, tau=10.0, v_th=20.0, r=1.0,
                v_reset=-5.0, g_tau=12.0, e_rev=0.0, izh_v_th=30.0, izh_r=1.0, e_ampa=0.0, e_gaba=-80.0,
                ampa_tau=5.0, gaba_tau=10.0)
    base.update(kw)
    keep = []
    m = O.OrcModel(**{k: v for k, v in base.items() if k not in "abcd"})
    for k in "abcd":
        if k in base:
            arr = np.ascontiguousarray(base[k], dtype=np.float32)
            keep.append(arr)
            setattr(m, k, O.ptr(arr, C.c_float))
    L_.orc_model_prepare(C.byref(m))
    m._keep = keep
    return m


def step(m, x, last, isyn=None):
    sp = np.zeros(max(1, m.n), np.int32)
    if x.dtype == np.float64:
        return L_.orc_update_f64(C.byref(m), O.ptr(x, C.c_double), O.ptr(last, C.c_int32),
                                 O.ptr(isyn, C.c_double) if isyn is not None else None, O.ptr(sp, C.c_int32)), sp
    return L_.orc_update_f32(C.byref(m), O.ptr(x, C.c_float), O.ptr(last, C.c_int32),
                             O.ptr(isyn, C.c_float) if isyn is not None else None, O.ptr(sp, C.c_int32)), sp


# ---------------------------------------------------------------- random numbers

def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 with 10 rounds
    assert list(O.philox([0, 0, 0, 0], [0, 0])) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert list(O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0])) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def _poisson_hash_py(neuron_index, time_step):
    # restated from numerical_accuracy_tests.py:118-128 (the reference's own Python check)
    x = (neuron_index * 747796405 + (time_step + 1) * 2891336453 + 0x9E3779B9) & 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 2246822519) & 0xFFFFFFFF
    x ^= x >> 13
    x = (x * 3266489917) & 0xFFFFFFFF
    x ^= x >> 16
    return (x & 0x00FFFFFF) / float(0x01000000)


def test_poisson_rate_hash_matches_reference_test():
    for neuron in (0, 1, 7, 12345):
        for t in range(0, 400, 3):
            assert L_.orc_hash_uniform01(neuron, t) == _poisson_hash_py(neuron, t)
    # the reference's expected PoissonRate sequence rule (numerical_accuracy_tests.py:139): 40 Hz, dt 1 ms
    fired = [t for t in range(200) if L_.orc_hash_uniform01(0, t) < np.float32(40.0 * 0.001)]
    assert fired == [t for t in range(200) if _poisson_hash_py(0, t) < 40.0 * 0.001]
    assert 2 <= len(fired) <= 20


def test_poisson_thresholds_and_counts():
    scipy_stats = pytest.importorskip("scipy.stats")
    for lam in (0.25, 1.0, 3.5):
        tab = np.zeros(64, np.uint32)
        kmax = L_.orc_poisson_thresholds(lam, O.ptr(tab, C.c_uint32), 64)
        cdf = scipy_stats.poisson.cdf(np.arange(kmax), lam)
        expect = np.minimum(np.floor(cdf * 2.0 ** 32), 2.0 ** 32 - 1)
        assert np.all(np.abs(tab[:kmax].astype(np.float64) - expect) <= 1.0)
        assert tab[kmax] == 0xFFFFFFFF
        counts = np.array([L_.orc_poisson_count(0x1234, 0x5678, i, t, O.ptr(tab, C.c_uint32), kmax)
                           for i in range(200) for t in range(100)])
        assert abs(counts.mean() - lam) < 4 * math.sqrt(lam / counts.size)
        hist = np.bincount(counts, minlength=kmax + 1)[:8].astype(np.float64)
        pmf = scipy_stats.poisson.pmf(np.arange(8), lam) * counts.size
        ok = pmf > 20
        chi2 = (((hist - pmf) ** 2) / pmf)[ok].sum()
        assert chi2 < scipy_stats.chi2.ppf(0.999, ok.sum() - 1)


# ---------------------------------------------------------------- neuron models (Euler)

def test_izhikevich_reference_known_answer():
    # numerical_accuracy_report.md: first run (2 internal substeps: I=0 then I=5, dt=1) gives
    # V=-65.03999328613281, U=-13.01200008392334 with max error 0 in float32
    m = model(model=O.IZH, v_rest=-65.0, a=[0.02], b=[0.2], c=[-65.0], d=[8.0])
    x = np.zeros(6, np.float32)
    last = np.zeros(1, np.int32)
    L_.orc_init_state(C.byref(m), O.ptr(x, C.c_float), O.ptr(last, C.c_int32))
    step(m, x, last)
    x[5] = 5.0
    step(m, x, last)
    assert float(x[0]) == -65.03999328613281
    assert float(x[1]) == -13.01200008392334


def _reference_lif(v_rest, tau, r, steps, initial_voltage, current=5.0, dt=1.0):
    # closed-form Euler recurrence of numerical_accuracy_tests.py:54-64 (2 substeps in the first run)
    v, out = initial_voltage, []
    for outer in range(steps):
        for sub in range(2 if outer == 0 else 1):
            cur = 0.0 if outer == 0 and sub == 0 else current
            v += ((r * cur / tau) + (v_rest - v) / tau) * dt
        out.append(v)
    return out


def test_voltage_jump_reference_known_answer():
    # numerical_accuracy_report.md: V=-5.0 -> -4.5 (no input) -> -3.55 (I=5), tolerance 2e-5
    m = model()
    x = np.array([-5.0, 0.0], np.float32)
    last = np.array([10000], np.int32)
    got = []
    for outer in range(24):
        for sub in range(2 if outer == 0 else 1):
            x[1] = 0.0 if outer == 0 and sub == 0 else 5.0
            step(m, x, last)
            if outer == 0 and sub == 0:
                assert float(x[0]) == pytest.approx(-4.5, abs=1e-6)
        got.append(float(x[0]))
    assert got[0] == pytest.approx(-3.55, abs=1e-6)
    assert np.max(np.abs(np.array(got) - _reference_lif(0.0, 10.0, 1.0, 24, -5.0))) <= 2e-5


def test_lif_exponential_decay_reference_known_answer():
    # numerical_accuracy_tests.py:67-69: V_rest -65, tau 20, R 1, I=5 through a current synapse
    m = model(model=O.LIF_EXP, v_rest=-65.0, tau=20.0, v_th=-20.0, v_reset=-65.0)
    x = np.array([-65.0, 0.0, 0.0], np.float32)
    last = np.array([10000], np.int32)
    got = []
    for outer in range(24):
        for sub in range(2 if outer == 0 else 1):
            x[2] = 0.0 if outer == 0 and sub == 0 else 5.0
            step(m, x, last)
        got.append(float(x[0]))
    assert np.max(np.abs(np.array(got) - _reference_lif(-65.0, 20.0, 1.0, 24, -65.0))) <= 2e-5


def test_voltage_jump_refractory_and_threshold():
    # spike when V > V_th (strict), reset, then t_ref steps with dV = 0 (Voltage_jump.cpp:57-87)
    m = model(v_th=-50.0, v_rest=-65.0, v_reset=-65.0, tau=20.0, t_ref=3, dt=0.1)
    x = np.array([-50.0, 0.0], np.float32)
    last = np.array([10000], np.int32)
    cnt, _ = step(m, x, last)  # decays below threshold: no spike at exactly V_th
    assert cnt == 0 and x[0] < -50.0
    x[0] = np.float32(-49.0)
    last[0] = 10000
    x[1] = 0.0
    # one Euler step from -49 moves toward rest but stays above -50 -> spike
    cnt, sp = step(m, x, last)
    assert cnt == 1 and x[0] == -65.0 and last[0] == 0
    x[1] = 10000.0  # huge drive (50 mV per step) is ignored while refractory
    # the gate is LastSpike > t_ref, tested before LastSpike += 1: t_ref + 1 frozen updates
    for k in range(4):
        cnt, _ = step(m, x, last)
        assert x[0] == -65.0 and cnt == 0
    assert last[0] == 4
    cnt, _ = step(m, x, last)
    assert cnt == 1  # integrates again; the drive crosses threshold at once


# ---------------------------------------------------------------- RK4 (no reference counterpart)

def test_lif_rk4_closed_form():
    # linear ODE: RK4 multiplies the distance to V_inf by 1 - h + h^2/2 - h^3/6 + h^4/24 per step
    tau, dt, I = 20.0, 0.1, 7.0
    m = model(integrator=O.RK4, v_rest=-65.0, tau=tau, v_th=100.0, v_reset=-65.0, dt=dt)
    h = float(np.float32(dt)) / tau  # the oracle's parameters are float32
    g = 1 - h + h * h / 2 - h ** 3 / 6 + h ** 4 / 24
    v_inf = -65.0 + I
    for dtype in (np.float64, np.float32):
        x = np.array([-70.0, I], dtype)
        last = np.array([10000], np.int32)
        for k in range(1, 501):
            step(m, x, last)
            expect = v_inf + (-70.0 - v_inf) * g ** k
            tol = 1e-11 if dtype == np.float64 else 1e-5 * abs(expect)
            assert abs(float(x[0]) - expect) <= tol, (dtype, k)


def test_izhikevich_rk4_fourth_order():
    # global error of RK4 falls 16x when dt halves (FP64 instantiation, subthreshold trajectory)
    def run(dt, T=20.0):
        m = model(model=O.IZH_CUR, integrator=O.RK4, dt=dt, v_rest=-70.0, a=[0.02], b=[0.2], c=[-65.0], d=[8.0])
        x = np.zeros(6, np.float64)
        x[0], x[1], x[5] = -75.0, -13.0, 1.0  # relaxes toward the stable rest point near -70
        last = np.array([10000], np.int32)
        for _ in range(int(round(T / dt))):
            cnt, _ = step(m, x, last)
            assert cnt == 0
        return x[0]
    ref = run(0.0078125)
    e1, e2 = abs(run(0.5) - ref), abs(run(0.25) - ref)  # powers of two: exact float32 dt and T
    assert 12.0 < e1 / e2 < 20.0


def test_fp32_step_error_against_fp64():
    # per-step FP32 error from identical state: the <= 1e-5 relative bound, measured
    rng = np.random.default_rng(7)
    n = 2000
    a = np.full(n, 0.02, np.float32); b = np.full(n, 0.2, np.float32)
    c = np.full(n, -65.0, np.float32); d = np.full(n, 8.0, np.float32)
    for integ in (O.EULER, O.RK4):
        m = model(model=O.IZH, integrator=integ, n=n, dt=0.1, v_rest=-65.0, a=a, b=b, c=c, d=d)
        x32 = np.zeros((n, 6), np.float32)
        x32[:, 0] = rng.uniform(-75, -50, n); x32[:, 1] = rng.uniform(-16, -10, n)
        x32[:, 2] = rng.uniform(0, 0.1, n); x32[:, 3] = rng.uniform(0, 0.1, n); x32[:, 5] = rng.uniform(0, 10, n)
        x64 = x32.astype(np.float64)
        l32 = np.full(n, 10000, np.int32); l64 = l32.copy()
        step(m, x32, l32)
        step(m, x64, l64)
        rel = np.abs(x32[:, :2] - x64[:, :2]) / np.maximum(np.abs(x64[:, :2]), 1.0)
        assert rel.max() <= 1e-5


# ---------------------------------------------------------------- delivery

def _small_sparse(n=600, k_in=60, seed=11):
    from neuronbridge_amd import workloads as W
    return W.sparse_lif(n=n, k_in=k_in, steps=300, seed=seed)


def test_push_ring_equals_pull_delivery_every_step():
    # orc_run checks each step that the push image ring[(t+d)%D] equals the pull delivery
    wl = _small_sparse()
    run = OracleRun(wl)
    counts, _ = run.run(300)
    assert counts.sum() > 50
    assert run.events[0] > 0


def test_ring_contents_brute_force():
    # after T steps the ring must hold, for each future step, the brute-force sum over the
    # recorded raster of q(w) for every synapse whose spike time + delay lands there
    wl = _small_sparse(n=400, k_in=40, seed=5)
    run = OracleRun(wl)
    T = 120
    counts, raster = run.run(T)
    n, D = wl.spec.n, wl.d_slots
    s = wl.syn
    scale = 2.0 ** run.shift
    q = np.rint(s.weight.astype(np.float32) * np.float32(scale)).astype(np.int64)
    q = np.where(s.type == 1, -q, q)
    ring = np.zeros((D, n), np.int64)
    for t in range(T - D, T):
        bits = np.unpackbits(raster[t].view(np.uint8), bitorder="little")[:n]
        for pre in np.nonzero(bits)[0]:
            for j in range(s.row_ptr[pre], s.row_ptr[pre + 1]):
                arr = t + int(s.delay[j])
                if arr >= T:
                    ring[arr % D, s.post[j]] += q[j]
    assert np.array_equal(ring, run.ring[:, 0, :].astype(np.int64))
    assert np.array_equal(counts, np.array([np.unpackbits(r.view(np.uint8)).sum() for r in raster]))


def test_delay_one_known_answer():
    # tests/native/dense_cuda_smoke.cpp:109-118: input fired at t=0 through a delay-1 synapse;
    # the target fires at t=1, not at t=0
    from neuronbridge_amd import _lib as LL
    from neuronbridge_amd.network import ModelSpec, Synapses
    from neuronbridge_amd.workloads import Workload
    spec = ModelSpec(model=LL.LIF_VJ, n=3, dt=0.1, t_ref=0, v_rest=-65.0, tau=20.0, v_th=-64.8, v_reset=-65.0)
    syn = Synapses(row_ptr=np.array([0, 1, 2, 2], np.int64), post=np.array([1, 2], np.int32),
                   delay=np.array([1, 1], np.uint8), type=np.zeros(2, np.uint8),
                   weight=np.array([12.0, 8.0], np.float32), max_delay=1)
    cur = np.zeros((4, 3), np.float32)
    cur[0, 0] = 1000.0
    wl = Workload("delay_kat", spec, syn=syn, cur=cur, steps=3)
    run = OracleRun(wl)
    _, raster = run.run(3)
    fired = [sorted(np.nonzero(np.unpackbits(r.view(np.uint8), bitorder="little")[:3])[0].tolist()) for r in raster]
    assert fired[0] == [0]
    assert fired[1] == [1]
    assert fired[2] == [2]


def test_fixed_point_vs_float_event_delivery_per_step():
    # device semantics (int32 fixed-point arrivals) vs the reference's per-event float adds:
    # from identical state, one step differs by <= 1e-5 relative
    wl = _small_sparse(n=800, k_in=80, seed=3)
    a = OracleRun(wl)
    b = OracleRun(wl)
    worst = 0.0
    for _ in range(60):
        a.run(1, O.ACC_FIXED, ring=False)
        b.run(1, O.ACC_FLOAT_EVENTS, ring=False)
        rel = np.abs(a.x[:, 0] - b.x[:, 0]) / np.maximum(np.abs(b.x[:, 0]), 1.0)
        worst = max(worst, float(rel.max()))
        b.x[:] = a.x
        b.last[:] = a.last
        b.hist[:] = a.hist
        b.hist_cnt[:] = a.hist_cnt
    assert worst <= 1e-5


# ---------------------------------------------------------------- STDP

def _stdp_continuous(pre_ticks, post_ticks, nticks, dt, ltp, tp, ltd, td, w0, wmax):
    """Exact-exponential pair STDP in float64 from its definition (traces as sums of kernels)."""
    w = w0
    ws = [w]
    pres, posts = [], []
    for t in range(nticks):
        if t in pre_ticks:
            apost = sum(math.exp(-(t - s) * dt / td) for s in posts)
            pres.append(t)
            w = min(max(w - ltd * apost, 0.0), wmax)
        if t in post_ticks:
            apre = sum(math.exp(-(t - s) * dt / tp) for s in pres)
            posts.append(t)
            w = min(max(w + ltp * apre, 0.0), wmax)
        ws.append(w)
    return np.array(ws)


def _replay(pre, post, nticks, **kw):
    p = np.zeros(nticks, np.uint8); p[list(pre)] = 1
    q = np.zeros(nticks, np.uint8); q[list(post)] = 1
    out = np.zeros(nticks + 1, np.float32)
    L_.orc_stdp_replay(kw["ltp"], kw["tp"], kw["ltd"], kw["td"], kw["dt"], kw["w0"], kw["wmax"],
                       O.ptr(p, C.c_uint8), O.ptr(q, C.c_uint8), nticks, O.ptr(out, C.c_float))
    return out


def test_stdp_isolated_pair_closed_form():
    kw = dict(ltp=0.01, tp=20.0, ltd=0.012, td=20.0, dt=0.5, w0=0.5, wmax=1.0)
    w = _replay([10], [30], 40, **kw)  # pre then post, 20 steps apart -> LTP = A+ exp(-10/20)
    assert w[-1] == pytest.approx(0.5 + 0.01 * math.exp(-10.0 / 20.0), abs=1e-6)
    w = _replay([30], [10], 40, **kw)  # post then pre -> LTD = A- exp(-10/20)
    assert w[-1] == pytest.approx(0.5 - 0.012 * math.exp(-10.0 / 20.0), abs=1e-6)
    w = _replay([5], [5], 10, **kw)  # same tick: pre side first (deliveries precede the update)
    assert w[-1] == pytest.approx(0.5 + 0.01, abs=1e-6)


def test_stdp_replay_scenario_and_clipping():
    # pair_stdp_gpu_trajectory.cpp: pre arrives at ticks 1 and 8 (spikes at 0 and 7, delay 1),
    # dt 1, reference STDP defaults (STDP.h:14-20), w 8 of 20; CPU tolerance 1e-6
    kw = dict(ltp=0.92, tp=16.8, ltd=0.53, td=33.1, dt=1.0, w0=8.0, wmax=20.0)
    pre, post = [1, 8], [1, 2, 8, 9, 12]
    got = _replay(pre, post, 16, **kw)
    ref = _stdp_continuous(set(pre), set(post), 16, 1.0, 0.92, 16.8, 0.53, 33.1, 8.0, 20.0)
    assert np.max(np.abs(got - ref)) <= 1e-5
    assert len(set(np.round(got, 6))) > 1
    kw2 = dict(kw, w0=19.9)
    hi = _replay([0, 1, 2, 3], [1, 2, 3, 4, 5], 8, **kw2)
    assert hi.max() == 20.0 and hi.min() >= 0.0
    kw3 = dict(kw, w0=0.1)
    lo = _replay([5, 6, 7], [0, 1, 2, 3, 4], 8, **kw3)
    assert lo.min() == 0.0


# ---- replay STDP (orc_replay_stdp): pinned to the oracle's own closed loop and to the
# single-synapse trajectory restatement above

def _stdp_run(T):
    from neuronbridge_amd import workloads as W
    wl = W.stdp_izh(n=800, k_in=50, steps=T, w_bg=2.0)
    run = OracleRun(wl, fixed_shift=20)
    counts, raster = run.run(T)
    return wl, run, counts, raster


def test_replay_stdp_equals_closed_loop_on_its_raster():
    from harness import raster_lists
    T = 600
    wl, run, counts, raster = _stdp_run(T)
    assert counts.sum() > 200
    lists = raster_lists(raster, wl.spec.n)
    fresh = OracleRun(wl, fixed_shift=20)
    upd = fresh.replay_stdp(lists)
    assert upd > 10000
    for name in ("w", "apre", "apost", "slast"):
        a, b = getattr(fresh, name), getattr(run, name)
        assert np.array_equal(a.view(np.uint32) if a.dtype == np.float32 else a,
                              b.view(np.uint32) if b.dtype == np.float32 else b), name
    # chained: the second call re-reads the last D - 1 steps of the first for their arrivals only
    D = wl.d_slots
    chained = OracleRun(wl, fixed_shift=20)
    t1 = 257
    u1 = chained.replay_stdp(lists[:t1], t0=0)
    u2 = chained.replay_stdp(lists[t1 - (D - 1):], t0=t1, lead=D - 1)
    assert u1 + u2 == upd
    assert np.array_equal(chained.w.view(np.uint32), run.w.view(np.uint32))
    assert np.array_equal(chained.slast, run.slast)
    # a call that leaves out the lead steps misses their arrivals: the pin is sensitive to it
    short = OracleRun(wl, fixed_shift=20)
    short.replay_stdp(lists[:t1], t0=0)
    short.replay_stdp(lists[t1:], t0=t1)
    assert not np.array_equal(short.w.view(np.uint32), run.w.view(np.uint32))


def test_replay_stdp_per_synapse_trajectories():
    """Each synapse's final weight equals the single-synapse schedule replay (orc_stdp_replay,
    pair_stdp_gpu_trajectory.cpp:52-66) on its arrival steps (pre spike + delay) and its
    target's spike steps."""
    from harness import raster_lists
    T = 400
    wl, run, counts, raster = _stdp_run(T)
    lists = raster_lists(raster, wl.spec.n)
    fresh = OracleRun(wl, fixed_shift=20)
    fresh.replay_stdp(lists)
    fired = np.zeros((T, wl.spec.n), np.uint8)
    for t, ids in enumerate(lists):
        fired[t, ids] = 1
    syn = wl.syn
    pre_of = np.repeat(np.arange(wl.spec.n), np.diff(syn.row_ptr))
    plastic = np.nonzero(syn.plastic)[0]
    rng = np.random.default_rng(5)
    sd = wl.stdp
    checked = moved = 0
    for s in rng.choice(plastic, 300, replace=False):
        d = int(syn.delay[s])
        pre_sched = np.zeros(T, np.uint8)
        pre_sched[d:] = fired[:T - d, pre_of[s]]
        post_sched = np.ascontiguousarray(fired[:, syn.post[s]])
        out = np.zeros(T + 1, np.float32)
        L_.orc_stdp_replay(sd["max_ltp"], sd["ltp_tau"], sd["max_ltd"], sd["ltd_tau"], wl.spec.dt,
                           float(syn.weight[s]), sd["w_max"], O.ptr(pre_sched, C.c_uint8),
                           O.ptr(post_sched, C.c_uint8), T, O.ptr(out, C.c_float))
        assert out[-1].view(np.uint32) == fresh.w[s].view(np.uint32), s
        checked += 1
        moved += out[-1] != syn.weight[s]
    assert checked == 300 and moved > 100


# ---------------------------------------------------------------- queue-parallel CPU baseline

@pytest.mark.parametrize("nq", [1, 3, 8])
@pytest.mark.parametrize("which", ["lif", "izh_cur"])
def test_queue_parallel_run_equals_float_event_run(nq, which):
    # the legacy OpenMP structure (a thread per queue of target neurons) delivers each target's
    # events in the single-thread order: identical results
    from neuronbridge_amd import workloads as W
    wl = W.sparse_lif(n=3000, k_in=80, integrator=L.RK4, steps=0, seed=3) if which == "lif" else \
        W.stdp_izh(n=3000, k_in=60, steps=0, w_bg=2.0)
    if which != "lif":
        wl.stdp = None
        wl.syn.plastic = None
    T = 250
    one = OracleRun(wl, fixed_shift=20)
    c1, _ = one.run(T, O.ACC_FLOAT_EVENTS, ring=False, raster=False)
    par = OracleRun(wl, fixed_shift=20)
    counts = np.zeros(T, np.int32)
    secs = C.c_double()
    st = L_.orc_run_queues(C.byref(par.net), nq, O.ptr(par.x, C.c_float), O.ptr(par.last, C.c_int32),
                           O.ptr(par.hist, C.c_int32), O.ptr(par.hist_cnt, C.c_int32), 0, T,
                           O.ptr(counts, C.c_int32), O.ptr(par.events, C.c_int64), C.byref(secs))
    assert st == 0
    assert c1.sum() > 100
    assert np.array_equal(counts, c1)
    assert np.array_equal(par.x.view(np.uint32), one.x.view(np.uint32))
    assert np.array_equal(par.last, one.last)
    assert par.events[0] == one.events[0]
    assert secs.value > 0.0
