"""GPU parity: the HIP library against the CPU oracle on identical seeded inputs.

Bar: bit-exact state (float32 bits), LastSpike, spike counts, rasters, delay-ring
contents, synaptic event counts and STDP weights/traces.
"""
import numpy as np
import pytest

import oracle as O
from harness import OracleRun
import neuronbridge_amd as nb
from neuronbridge_amd import _lib as L
from neuronbridge_amd import workloads as W
from neuronbridge_amd.network import ModelSpec, Synapses
from neuronbridge_amd.workloads import Workload

pytestmark = pytest.mark.gpu

ENGINES = {"kernels": L.ENGINE_KERNELS, "graph": L.ENGINE_GRAPH, "persistent": L.ENGINE_PERSISTENT}


def assert_bits_equal(a, b, what):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.shape == b.shape, what
    if a.dtype == np.float32:
        bad = np.nonzero(a.view(np.uint32) != b.view(np.uint32))
    else:
        bad = np.nonzero(a != b)
    assert bad[0].size == 0, f"{what}: {bad[0].size} mismatches, first at {[x[:3] for x in bad]}: " \
                             f"{a[tuple(x[:1] for x in bad)]} vs {b[tuple(x[:1] for x in bad)]}"


def compare(net, run, wl, counts_o, raster_o, ring=True):
    x, last = net.get_state()
    assert_bits_equal(x, run.x, "state")
    assert_bits_equal(last, run.last, "LastSpike")
    assert_bits_equal(net.counts(), counts_o, "spike counts")
    assert_bits_equal(net.raster(), raster_o, "raster")
    assert net.events() == int(run.events[0]), "synaptic events"
    if ring:
        assert_bits_equal(net.ring(), run.ring_as_device(wl.spec.channels), "delay ring")


def small_sparse(model, integrator, n=2000, k_in=100, seed=21):
    wl = W.sparse_lif(n=n, k_in=k_in, integrator=integrator, steps=400, seed=seed)
    if model == L.LIF_VJ:
        return wl
    rng = np.random.default_rng(seed)
    n_exc = wl.extra["n_exc"]
    if model == L.LIF_EXP:
        wl.spec = ModelSpec(model=L.LIF_EXP, n=n, dt=0.1, integrator=integrator, t_ref=20, v_rest=-65.0, tau=20.0,
                            v_th=-50.0, r=1.0, v_reset=-65.0, g_tau=5.0, e_rev=0.0)
        wl.syn.weight = np.where(wl.syn.type == 0, 0.02, 0.05).astype(np.float32)
        wl.w_bg = 0.03
        x0 = np.zeros((n, 3), np.float32)
        x0[:, 0] = wl.x0[:, 0]
        wl.x0 = x0
    else:  # conductance Izhikevich
        a, b, c, d = W.izh_params(n, n_exc, rng)
        wl.spec = ModelSpec(model=L.IZH, n=n, dt=0.1, integrator=integrator, v_rest=-65.0, a=a, b=b, c=c, d=d)
        wl.syn.weight = np.where(wl.syn.type == 0, 0.05, 0.1).astype(np.float32)
        wl.w_bg = 0.2
        x0 = np.zeros((n, 6), np.float32)
        x0[:, 0] = rng.uniform(-70, -55, n)
        x0[:, 1] = b * x0[:, 0]
        wl.x0 = x0
    return wl


@pytest.mark.parametrize("engine", list(ENGINES))
@pytest.mark.parametrize("integrator", [L.EULER, L.RK4], ids=["euler", "rk4"])
def test_sparse_lif_closed_loop_bitexact(gpu, engine, integrator):
    wl = small_sparse(L.LIF_VJ, integrator)
    T = 400
    net = wl.make_net()
    net.record(0, T, counts=True, raster=True)
    net.run(T, ENGINES[engine])
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    counts_o, raster_o = run.run(T)
    assert counts_o.sum() > 100, "workload too quiet to test delivery"
    compare(net, run, wl, counts_o, raster_o)


@pytest.mark.parametrize("model,integrator", [(L.LIF_EXP, L.EULER), (L.LIF_EXP, L.RK4), (L.IZH, L.RK4),
                                               (L.IZH, L.EULER)],
                         ids=["lif_exp_euler", "lif_exp_rk4", "izh_cond_rk4", "izh_cond_euler"])
@pytest.mark.parametrize("engine", ["kernels", "graph", "persistent"])
def test_other_models_closed_loop_bitexact(gpu, model, integrator, engine):
    # the conductance models with both integrators, the reference's (Euler) included, on every engine
    wl = small_sparse(model, integrator)
    T = 300
    net = wl.make_net()
    net.record(0, T, counts=True, raster=True)
    net.run(T, ENGINES[engine])
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    counts_o, raster_o = run.run(T)
    assert counts_o.sum() > 20
    compare(net, run, wl, counts_o, raster_o)


def test_engines_interoperate(gpu):
    # state, ring and firing history carry over between engines mid-run
    wl = small_sparse(L.LIF_VJ, L.RK4, n=3000, k_in=120, seed=4)
    net = wl.make_net()
    net.record(0, 330, counts=True, raster=True)
    net.run(100, L.ENGINE_GRAPH)
    net.run(130, L.ENGINE_PERSISTENT)
    net.run(37, L.ENGINE_KERNELS)
    net.run(63, L.ENGINE_PERSISTENT)
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    counts_o, raster_o = run.run(330)
    compare(net, run, wl, counts_o, raster_o)
    assert_history(net, raster_o, wl.spec.n)


@pytest.mark.parametrize("integrator", [L.EULER, L.RK4], ids=["euler", "rk4"])
def test_persistent_burst_steps_bitexact(gpu, integrator):
    # every neuron starts just under threshold, so the first steps fire thousands of neurons at
    # once: more spikes than one workgroup's step record holds and than the engine keeps in LDS
    wl = W.sparse_lif(n=20_000, k_in=100, integrator=integrator, steps=60, seed=17)
    wl.x0[:, 0] = np.linspace(-50.02, -50.0, wl.spec.n, endpoint=False, dtype=np.float32)
    net = wl.make_net()
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    counts_o, raster_o = run.run(60)
    assert counts_o[:2].min() > 2048, "no burst step"
    net.record(0, 60, counts=True, raster=True)
    net.run(6, L.ENGINE_PERSISTENT)
    bits = np.unpackbits(raster_o.view(np.uint8), axis=1, bitorder="little")[:, :wl.spec.n]
    for t in range(6):  # the first launch's firing history holds the burst steps
        assert_bits_equal(np.sort(net.history(t)), np.nonzero(bits[t])[0].astype(np.int32), f"history of step {t}")
    net.run(31, L.ENGINE_PERSISTENT)
    net.run(23, L.ENGINE_PERSISTENT)
    compare(net, run, wl, counts_o, raster_o)
    assert_history(net, raster_o, wl.spec.n)
    net.close()


def test_persistent_residency_failure_takes_no_step(gpu, monkeypatch):
    # a launch whose workgroups cannot all be resident (simulated from the host: the agreement waits
    # for one workgroup more than the grid has, with a short wait) takes no step and reports an
    # error; after a reset the engine runs bit-exact again
    wl = small_sparse(L.LIF_VJ, L.RK4)
    net = wl.make_net()
    x_init, l_init = net.get_state()
    net.run(40, L.ENGINE_PERSISTENT)
    x0, l0 = net.get_state()
    monkeypatch.setenv("NBX_PERSISTENT_TEST_ABSENT", "1")
    with pytest.raises(Exception, match="resident"):
        net.run(10, L.ENGINE_PERSISTENT)
    x1, l1 = net.get_state()
    assert_bits_equal(x1, x0, "state after the failed launch")
    assert_bits_equal(l1, l0, "LastSpike after the failed launch")
    monkeypatch.delenv("NBX_PERSISTENT_TEST_ABSENT")
    with pytest.raises(Exception, match="resident"):  # the error stays until a reset
        net.run(5, L.ENGINE_PERSISTENT)
    net.reset_state()
    net.set_state(x_init, l_init)
    T = 300
    net.record(0, T, counts=True, raster=True)
    net.run(T, L.ENGINE_PERSISTENT)
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    counts_o, raster_o = run.run(T)
    compare(net, run, wl, counts_o, raster_o)
    net.close()


def assert_history(net, raster_o, n):
    """Every step the firing-history ring still holds lists exactly the raster's spikes."""
    slots = net.ring().shape[0]
    T = raster_o.shape[0]
    bits = np.unpackbits(raster_o.view(np.uint8), axis=1, bitorder="little")[:, :n]
    for t in range(T - slots + 2, T):
        assert_bits_equal(np.sort(net.history(t)), np.nonzero(bits[t])[0].astype(np.int32), f"history of step {t}")


@pytest.mark.parametrize("engine", ["graph", "persistent", "mixed"])
def test_dense_izhikevich_current_config(gpu, engine):
    # dense all-to-all current-based Izhikevich at the configuration's full size (N=1000); the
    # persistent engine runs it in one workgroup, "mixed" alternates all three engines
    T = 1500
    wl = W.dense_izh(n=1000, steps=T, noise_scale=10 ** 0.5)
    net = wl.make_net()
    net.record(0, T, counts=True, raster=True)
    if engine == "mixed":
        for k, e in ((411, L.ENGINE_PERSISTENT), (300, L.ENGINE_GRAPH), (1, L.ENGINE_PERSISTENT),
                     (250, L.ENGINE_KERNELS), (T - 962, L.ENGINE_PERSISTENT)):
            net.run(k, e)
    else:
        assert net.pick_engine() == L.ENGINE_PERSISTENT
        net.run(T, ENGINES[engine])
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    counts_o, raster_o = run.run(T)
    assert counts_o.sum() > 100
    compare(net, run, wl, counts_o, raster_o)


def coincident_updates(wl, raster):
    """Plastic synapses whose target fired at t - 1 while a presynaptic spike arrives at t: the
    pairs for which one launch applies the post-side update of t - 1 and the pre-side update of t."""
    fired = np.unpackbits(raster.view(np.uint8), axis=1, bitorder="little")[:, :wl.spec.n].astype(bool)
    syn = wl.syn
    pre = np.repeat(np.arange(wl.spec.n), np.diff(syn.row_ptr))
    plastic = syn.plastic == 1
    total = 0
    for d in np.unique(syn.delay[plastic]):
        m = plastic & (syn.delay == d)
        a, b = pre[m], syn.post[m]
        # target fired at t - 1 (rows 0 .. T-3) and the source at t - d (rows 1 - d .. T-1-d)
        for t in range(max(1, int(d)), fired.shape[0]):
            total += int(np.count_nonzero(fired[t - 1, b] & fired[t - int(d), a]))
    return total


@pytest.mark.parametrize("split", [False, True], ids=["one-run", "three-runs"])
@pytest.mark.parametrize("engine", ["kernels", "graph"])
def test_stdp_weights_bitexact(gpu, engine, split):
    wl = W.stdp_izh(n=1500, k_in=60, steps=1200, w_bg=2.0)
    T = 1200
    net = wl.make_net()
    net.record(0, T, counts=True, raster=True)
    if split:  # each run opens with a plain pull and closes with its own post-side launch
        for k in (333, 64, T - 397):
            net.run(k, ENGINES[engine])
    else:
        net.run(T, ENGINES[engine])
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    counts_o, raster_o = run.run(T)
    assert counts_o.sum() > 200, "too few spikes for plasticity"
    assert coincident_updates(wl, raster_o) > 50, "no synapse takes both updates in one step"
    compare(net, run, wl, counts_o, raster_o, ring=False)
    w = net.weights()
    assert_bits_equal(w, run.w, "STDP weights")
    apre, apost, last = net.stdp_state()
    assert_bits_equal(apre, run.apre, "Apre")
    assert_bits_equal(apost, run.apost, "Apost")
    assert_bits_equal(last, run.slast, "LastUpdate")
    assert not np.array_equal(w, wl.syn.weight), "weights never changed"
    pw = w[wl.syn.plastic == 1]
    assert pw.min() >= 0.0 and pw.max() <= 1.0  # plastic weights stay clipped to [0, w_max]
    assert np.array_equal(w[wl.syn.plastic == 0], wl.syn.weight[wl.syn.plastic == 0])


def test_per_phase_delay_known_answer(gpu):
    # dense_cuda_smoke.cpp:109-118 through the per-phase calls: an input spike injected at
    # t=0 through a delay-1 synapse makes the target fire at t=1, not t=0
    spec = ModelSpec(model=L.LIF_VJ, n=3, dt=0.1, t_ref=0, v_rest=-65.0, tau=20.0, v_th=-64.8, v_reset=-65.0)
    syn = Synapses(row_ptr=np.array([0, 1, 2, 2], np.int64), post=np.array([1, 2], np.int32),
                   delay=np.array([1, 1], np.uint8), type=np.zeros(2, np.uint8),
                   weight=np.array([12.0, 8.0], np.float32), max_delay=1)
    net = Workload("kat", spec, syn=syn).make_net()
    fired = []
    for t in range(3):
        net.begin_step(t)
        net.run_propagation()
        ids = net.neuron_step()
        if t == 0:
            net.inject_spikes([0])
        net.apply_post_learning()
        net.commit_firings()
        fired.append(sorted(net.history(t).tolist()))
    assert fired == [[0], [1], [2]]
    with pytest.raises(L.NbxError):
        net.begin_step(7)  # out of order: the reference's BeginStep contract


def replay_expect(wl, fixed_shift, spikes):
    """Brute force: every synapse of every injected spike adds its fixed-point weight at step
    k + delay - 1; inhibitory synapses go to channel 1 in the two-channel models (conductance
    Izhikevich, double-exponential LIF, trigger relay), every type to its own channel in the
    triple-exponential LIF, and enter negated in the one-channel LIF and Poisson-rate models."""
    s = wl.syn
    K = len(spikes)
    model = wl.spec.model
    q = np.rint(s.weight * np.float32(2.0 ** fixed_shift)).astype(np.int64)
    if model in (L.LIF_VJ, L.LIF_EXP, L.POISSON):
        q = np.where(s.type == 1, -q, q)
    nch = wl.spec.channels
    if nch == 3:
        ch = s.type.astype(np.int64)
    elif nch == 2:
        ch = (s.type == 1).astype(np.int64)
    else:
        ch = np.zeros(s.type.shape, np.int64)
    expect = np.zeros((K, nch, wl.spec.n), np.int64)
    events = 0
    for k, ids in enumerate(spikes):
        if len(ids) == 0:
            continue
        sl = np.concatenate([np.arange(s.row_ptr[p], s.row_ptr[p + 1]) for p in ids])
        events += sl.size
        ko = k + s.delay[sl].astype(np.int64) - 1
        ok = ko < K
        np.add.at(expect, (ko[ok], ch[sl][ok], s.post[sl][ok]), q[sl][ok])
    return expect, events


# Replay launch shapes forced through the library's launch-time knobs: (merge factor, window,
# time groups, staged spikes, slot layout, streamed slot kernel, packed 6-byte slot entries);
# None = the geometry the library picks. "c64" makes most windows hold more spikes than one
# staging buffer (the chunked path); the streamed kernel has no windows or staging (its ring holds
# as many steps as fit; W caps its super step).
REPLAY_KNOBS = ("NBX_REPLAY_M", "NBX_REPLAY_W", "NBX_REPLAY_G", "NBX_REPLAY_C", "NBX_REPLAY_SLOTS",
                "NBX_REPLAY_STREAM", "NBX_REPLAY_PACKED")
REPLAY_SHAPES = {
    "auto": None,
    "ring-m2w8g2": ("2", "8", "2", "", "0", "0", "0"),
    "ring-m1w8g3": ("1", "8", "3", "", "0", "0", "0"),
    "ring-m2w16g5": ("2", "16", "5", "", "0", "0", "0"),
    "ring-m1w12g2c64": ("1", "12", "2", "64", "0", "0", "0"),
    "slots-m2w8g2": ("2", "8", "2", "", "1", "0", "0"),
    "slots-m1w12g3c64": ("1", "12", "3", "64", "1", "0", "0"),
    "stream-m2g2": ("2", "", "2", "", "1", "1", "0"),
    "stream-m3g3": ("3", "", "3", "", "1", "1", "0"),
    "stream-m1g7w1": ("1", "1", "7", "", "1", "1", "0"),
    "packed-m3g3": ("3", "", "3", "", "1", "1", "1"),
    "packed-m2g2w2": ("2", "2", "2", "", "1", "1", "1"),
    "packed-m1g5": ("1", "", "5", "", "1", "1", "1"),
}


def set_replay_shape(monkeypatch, shape):
    if shape:
        for k, v in zip(REPLAY_KNOBS, shape):
            monkeypatch.setenv(k, v)


def random_spikes(rng, n, K, lo, hi):
    return [np.sort(rng.choice(n, size=int(rng.integers(lo, hi)), replace=False)).astype(np.int32) for _ in range(K)]


@pytest.mark.parametrize("shape", list(REPLAY_SHAPES.values()), ids=list(REPLAY_SHAPES))
def test_replay_delivery_brute_force(gpu, shape, monkeypatch):
    set_replay_shape(monkeypatch, shape)
    wl = small_sparse(L.LIF_VJ, L.EULER, n=2500, k_in=80, seed=9)
    net = wl.make_net()
    rng = np.random.default_rng(1)
    K = 150  # spans several replay windows and, with time groups, their halos
    spikes = random_spikes(rng, wl.spec.n, K, 0, 60)
    spikes[7] = spikes[7][:0]  # an empty step
    if shape:
        expect_path = L.REPLAY_OFFSETS if shape[4] == "0" else (L.REPLAY_STREAM if shape[5] == "1" else L.REPLAY_SLOTS)
        assert net.replay_path() == expect_path
    e0 = net.events()
    out = net.replay(spikes)
    expect, events = replay_expect(wl, net.fixed_shift, spikes)
    assert_bits_equal(out.astype(np.int64), expect, "replay arrivals")
    assert net.events() - e0 == events


STREAM_VARIANTS = {"window": ("2", "8", "0", "0"), "stream": ("2", "", "1", "0"), "packed": ("3", "", "1", "1")}


def stream_shape(variant, g):
    m, w, stream, packed = STREAM_VARIANTS[variant]
    return (m, w, g, "", "1", stream, packed)


@pytest.mark.parametrize("variant", list(STREAM_VARIANTS))
def test_replay_dense_steps(gpu, variant, monkeypatch):
    # 600-1300 spikes per step: the streamed kernel takes one input step per super step and
    # several chunks per wave; the window kernel stages in several chunks
    set_replay_shape(monkeypatch, stream_shape(variant, "3"))
    wl = small_sparse(L.LIF_VJ, L.EULER, n=2500, k_in=80, seed=11)
    net = wl.make_net()
    rng = np.random.default_rng(7)
    K = 40
    spikes = random_spikes(rng, wl.spec.n, K, 600, 1300)
    e0 = net.events()
    out = net.replay(spikes)
    expect, events = replay_expect(wl, net.fixed_shift, spikes)
    assert_bits_equal(out.astype(np.int64), expect, "replay arrivals of dense steps")
    assert net.events() - e0 == events


@pytest.mark.parametrize("variant", list(STREAM_VARIANTS))
def test_replay_slot_overflow(gpu, variant, monkeypatch):
    # 640 neurons with in-degree 400: two or three 32-neuron blocks receive ~40-60 synapses per
    # presynaptic neuron, so most slots continue in the overflow list
    set_replay_shape(monkeypatch, stream_shape(variant, "2"))
    wl = small_sparse(L.LIF_VJ, L.EULER, n=640, k_in=400, seed=4)
    net = wl.make_net()
    rng = np.random.default_rng(2)
    K = 60
    spikes = random_spikes(rng, wl.spec.n, K, 0, 40)
    e0 = net.events()
    out = net.replay(spikes)
    expect, events = replay_expect(wl, net.fixed_shift, spikes)
    assert_bits_equal(out.astype(np.int64), expect, "replay arrivals with slot overflow")
    assert net.events() - e0 == events


@pytest.mark.parametrize("slots", ["0", "1", "stream", "packed"], ids=["ring", "slots", "stream", "packed"])
def test_replay_two_channels(gpu, slots, monkeypatch):
    # conductance Izhikevich: inhibitory arrivals go to their own channel
    shapes = {"stream": ("2", "", "2", "", "1", "1", "0"), "packed": ("2", "", "2", "", "1", "1", "1")}
    set_replay_shape(monkeypatch, shapes.get(slots, ("2", "8", "2", "", slots, "0", "0")))
    wl = small_sparse(L.IZH, L.EULER, n=2500, k_in=80, seed=6)
    net = wl.make_net()
    rng = np.random.default_rng(3)
    K = 70
    spikes = random_spikes(rng, wl.spec.n, K, 0, 50)
    e0 = net.events()
    out = net.replay(spikes)
    expect, events = replay_expect(wl, net.fixed_shift, spikes)
    assert_bits_equal(out.astype(np.int64), expect, "two-channel replay arrivals")
    assert net.events() - e0 == events


@pytest.mark.parametrize("slots", ["auto", "0", "window", "stream"], ids=["auto", "ring", "window", "stream"])
def test_replay_full_size(gpu, slots, monkeypatch):
    # configuration-size network (1e5 neurons, 1e8 synapses), 96 steps of 100-220 spikes; "auto"
    # is the streamed kernel on packed slots of three blocks, "stream" on 8-byte slots of two
    if slots == "window":
        monkeypatch.setenv("NBX_REPLAY_STREAM", "0")
    elif slots == "stream":
        monkeypatch.setenv("NBX_REPLAY_PACKED", "0")
    elif slots != "auto":
        monkeypatch.setenv("NBX_REPLAY_SLOTS", slots)
    wl = W.sparse_lif()
    net = wl.make_net()
    expect_path = {"auto": L.REPLAY_STREAM, "0": L.REPLAY_OFFSETS, "window": L.REPLAY_SLOTS,
                   "stream": L.REPLAY_STREAM}[slots]
    assert net.replay_path() == expect_path
    rng = np.random.default_rng(5)
    K = 96
    spikes = random_spikes(rng, wl.spec.n, K, 100, 220)
    e0 = net.events()
    out = net.replay(spikes)
    expect, events = replay_expect(wl, net.fixed_shift, spikes)
    assert_bits_equal(out.astype(np.int64), expect, "full-size replay arrivals")
    assert net.events() - e0 == events


def test_full_size_sparse_properties(gpu):
    # configuration-size network (1e5 neurons, 1e8 synapses): engines agree bit for bit,
    # event count = sum over spikes of out-degree, raster popcount = spike count
    wl = W.sparse_lif()
    T = 150
    a = wl.make_net()
    a.record(0, T, counts=True, raster=True)
    a.run(T, L.ENGINE_PERSISTENT)
    b = wl.make_net()
    b.record(0, T, counts=True, raster=True)
    b.run(T, L.ENGINE_GRAPH)
    xa, la = a.get_state()
    xb, lb = b.get_state()
    assert_bits_equal(xa, xb, "state persistent vs graph")
    assert_bits_equal(la, lb, "LastSpike persistent vs graph")
    ra, rb = a.raster(), b.raster()
    assert_bits_equal(ra, rb, "raster persistent vs graph")
    assert_bits_equal(a.ring(), b.ring(), "ring persistent vs graph")
    counts = a.counts()
    assert np.array_equal(counts, [int(np.unpackbits(r.view(np.uint8)).sum()) for r in ra])
    deg = np.diff(wl.syn.row_ptr)
    bits = np.unpackbits(ra.view(np.uint8), axis=1, bitorder="little")[:, :wl.spec.n]
    assert a.events() == int((bits.astype(np.int64) * deg[None, :]).sum())
    assert a.events() == b.events()
    assert counts.sum() > 0
