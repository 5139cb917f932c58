"""Current synapses (synapse type 3) and injected currents on the GPU against the oracle.

Semantics (the reference's InputCurrent -> PropogatedCurrent -> ProcessCurrent chain,
Event/src/Current/InputCurrent.cpp:13-24, PropogatedCurrent.cpp:15-25,
TimeDrivenLIF_Voltage_jump.cpp:117-147, Current/src/CurrentSynapes.cpp): a current injected at a
source for step t reaches each of its current synapses at t + delay and sets that synapse's slot
(no weight applied); the target's I_ext becomes the float sum of its slots in declaration order.
The expected I_ext of every step is resolved here from those rules and handed to the oracle as
stateful current inputs; the GPU runs the whole network bit for bit against it.
"""
import numpy as np
import pytest

from harness import OracleRun
from neuronbridge_amd import _lib as L
from test_gpu_parity import ENGINES, compare, small_sparse

pytestmark = pytest.mark.gpu


def expected_inputs(n, src, dst, delay, events, T):
    """Per step the (target, 3, new I_ext) inputs of the touched targets."""
    order = [[] for _ in range(n)]
    for s, tg in enumerate(dst):
        order[tg].append(s)
    arrivals = [[] for _ in range(T)]
    for t, srcs, vals in events:  # injection order
        for so, v in zip(srcs, vals):
            for s in np.nonzero(src == so)[0]:
                if t + int(delay[s]) < T:
                    arrivals[t + int(delay[s])].append((int(s), np.float32(v)))
    slot = np.zeros(len(src), np.float32)
    out = [[] for _ in range(T)]
    for t in range(T):
        touched = set()
        for s, v in arrivals[t]:
            slot[s] = v
            touched.add(int(dst[s]))
        for tg in sorted(touched):
            c = np.float32(0.0)
            for s in order[tg]:
                c = np.float32(c + slot[s])
            out[t].append((tg, 3, float(c)))
    return out


def setup(seed=11):
    wl = small_sparse(L.LIF_VJ, L.RK4)
    n = wl.spec.n
    rng = np.random.default_rng(seed)
    drivers = rng.choice(n, 12, replace=False)
    m = 400
    src = rng.choice(drivers, m).astype(np.int32)
    dst = rng.integers(0, n, m).astype(np.int32)
    dst[:40] = dst[40:80]  # targets with several slots
    delay = rng.integers(1, 11, m).astype(np.uint8)
    events = []
    for t in (0, 3, 17, 18, 60, 61, 140):
        k = int(rng.integers(2, 8))
        so = rng.choice(drivers, k, replace=False)
        events.append((t, so, rng.uniform(-2.0, 6.0, k).astype(np.float32)))
    events.append((60, [drivers[0]], np.array([9.0], np.float32)))  # same step again: a later current wins
    return wl, src, dst, delay, events


@pytest.mark.parametrize("engine", ["graph", "persistent", "per_phase"])
def test_current_synapses_bitexact(gpu, engine):
    wl, src, dst, delay, events = setup()
    T = 250
    net = wl.make_net()
    net.set_current_synapses(src, dst, delay)
    net.record(0, T, counts=True, raster=True)
    by_t = {}
    for t, so, v in events:
        by_t.setdefault(t, []).append((so, v))
    if engine == "per_phase":
        for t in range(T):
            for so, v in by_t.get(t, []):
                net.inject_currents(so, v, t)
            net.begin_step(t)
            net.run_propagation()
            net.neuron_step()
            net.commit_firings()
    else:
        eng = ENGINES[engine]
        done = 0
        for t in sorted(by_t) + [T]:  # inject at each event's step, running up to it first
            if t > done:
                net.run(t - done, eng)
                done = t
            for so, v in by_t.get(t, []):
                net.inject_currents(so, v, t)
    run = OracleRun(wl, fixed_shift=net.fixed_shift)
    run.set_inputs(expected_inputs(wl.spec.n, src, dst, delay, events, T), 0)
    counts_o, raster_o = run.run(T)
    compare(net, run, wl, counts_o, raster_o)
    x, _ = net.get_state()
    assert (x[:, 1] != 0).sum() > 20, "too few currents arrived"
    net.close()


def test_current_slots_known_answer(gpu):
    # two current synapses into neuron 5 (slots 0, 1) and one into neuron 6, from sources 0 and 1
    wl = small_sparse(L.LIF_VJ, L.EULER, n=300, k_in=10)
    net = wl.make_net()
    net.set_current_synapses([0, 1, 0], [5, 5, 6], [2, 3, 1])
    net.inject_currents([0], [0.5], 0)     # slot 0 of 5 at step 2, neuron 6 at step 1
    net.inject_currents([1], [0.25], 0)    # slot 1 of 5 at step 3
    net.inject_currents([0], [-1.0], 2)    # slot 0 of 5 at step 4, neuron 6 at step 3
    got = []
    for t in range(6):
        net.run(1, L.ENGINE_GRAPH)
        x, _ = net.get_state()
        got.append((float(x[5, 1]), float(x[6, 1])))
    assert got == [(0.0, 0.0), (0.0, 0.5), (0.5, 0.5), (0.75, -1.0), (-0.75, -1.0), (-0.75, -1.0)]
    net.reset_state()
    net.run(5, L.ENGINE_GRAPH)
    x, _ = net.get_state()
    assert x[5, 1] == 0.0 and x[6, 1] == 0.0  # ResetState clears slots and pending currents
    with pytest.raises(L.NbxError):
        net.inject_currents([0], [1.0], 2)      # before the current step
    with pytest.raises(L.NbxError):
        net.set_current_synapses([0], [300], [1])
    with pytest.raises(L.NbxError):
        net.set_current_synapses([0], [1], [0])
