"""Pins of oracle functions that the reference's known answers leave open.

* StableUnitRandom (CerebullarLearningRule.cpp:10-24) against vectors written by the reference's
  own function, compiled from its source (tests/golden/make_stable_unit_random.py), and live
  against that build when it is present.
* The conductance terms of every conductance model, by closed forms that do not restate the
  oracle's arithmetic: with the conductances held (decay factor 1) and a constant current, each
  model's membrane equation is linear in V, so Euler and RK4 contract towards the fixed point
  V_inf = (sum of g_k E_k + R I + V_rest) / (1 + sum of g_k) by a known factor per step; the
  Izhikevich conductance current vanishes at each reversal potential, changes sign across it and
  adds linearly; conductances decay by exp(-n dt / tau) after n steps.

Reference paths are relative to src/native/legacy/source_file_realtime_v1_async/NeuralModel/src/TimeDriven/.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

import oracle as O
from test_oracle import model, step

L_ = O.lib()
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "stable_unit_random.npz")
REF_SUR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                       "libref_sur.so")
ALL_COND = O.COND_EXC | O.COND_INH | O.COND_NMDA


# ---------------------------------------------------------------- StableUnitRandom

def test_stable_unit_random_matches_reference_vectors():
    g = np.load(GOLDEN)
    got = np.array([L_.orc_stable_unit_random(int(s), int(r), int(y), int(t))
                    for s, r, y, t in zip(g["seed"], g["rule"], g["synapse"], g["step"])], np.float32)
    assert got.size == g["value_bits"].size >= 4096
    assert np.array_equal(got.view(np.uint32), g["value_bits"])


@pytest.mark.skipif(not os.path.exists(REF_SUR), reason="oracle/_ref not built (needs the reference tree)")
def test_stable_unit_random_matches_reference_build_live():
    lib = C.CDLL(REF_SUR)
    f = lib.ref_stable_unit_random
    f.restype = C.c_float
    f.argtypes = [C.c_uint32, C.c_int32, C.c_int32, C.c_int32]
    rng = np.random.default_rng(7)
    for _ in range(20000):
        a = int(rng.integers(0, 2 ** 32))
        b, c, d = (int(v) for v in rng.integers(-2 ** 31, 2 ** 31, 3))
        assert L_.orc_stable_unit_random(a, b, c, d) == f(a, b, c, d)


# ---------------------------------------------------------------- conductance closed forms

def _rk4_factor(h):
    return 1.0 - h + h * h / 2.0 - h ** 3 / 6.0 + h ** 4 / 24.0


def _run(m, x, steps, dtype):
    last = np.array([10000], np.int32)
    out = []
    for _ in range(steps):
        step(m, x, last)
        out.append(float(x[0]))
    return np.array(out)


def _closed(v0, v_inf, factor, steps):
    k = np.arange(1, steps + 1)
    return v_inf + (v0 - v_inf) * factor ** k


@pytest.mark.parametrize("integ", [O.EULER, O.RK4])
def test_lif_exp_conductance_closed_form(integ):
    # TimeDrivenLIF_Exponential_Decay.cpp:62-76: tau dV/dt = R I + g (E - V) + (V_rest - V); g_tau = 0
    # gives the single-entry lookup factor exp(-dt * 0) = 1, so g is held.
    dt, tau, r, E, v_rest, I, g = 0.125, 20.0, 1.5, 0.0, -65.0, 4.0, 0.35
    m = model(model=O.LIF_EXP, integrator=integ, dt=dt, tau=tau, r=r, e_rev=E, v_rest=v_rest, v_th=100.0,
              g_tau=0.0)
    assert m.g_decay == 1.0
    x = np.array([-70.0, g, I], np.float64)
    got = _run(m, x, 200, np.float64)
    v_inf = (r * I + g * E + v_rest) / (1.0 + g)
    h = (1.0 + g) * dt / tau
    fac = 1.0 - h if integ == O.EULER else _rk4_factor(h)
    np.testing.assert_allclose(got, _closed(-70.0, v_inf, fac, 200), rtol=1e-12)
    # float32: the same trajectory within rounding
    x32 = np.array([-70.0, g, I], np.float32)
    np.testing.assert_allclose(_run(m, x32, 200, np.float32), _closed(-70.0, v_inf, fac, 200), rtol=2e-6)


@pytest.mark.parametrize("integ", [O.EULER, O.RK4])
def test_lif_exp_double_conductance_closed_form(integ):
    # TimeDrivenLIF_Exponential_double.cpp:61-74: tau dV/dt = R I + ge (Ee - V) + gi (Ei - V) + (V_rest - V)
    dt, tau, r, Ee, Ei, v_rest, I, ge, gi = 0.125, 15.0, 2.0, 0.0, -80.0, -65.0, 3.0, 0.2, 0.45
    m = model(model=O.LIF_EXP2, integrator=integ, dt=dt, tau=tau, r=r, e_ampa=Ee, e_gaba=Ei, v_rest=v_rest,
              v_th=100.0, ampa_tau=0.0, gaba_tau=0.0)
    m.cond = O.COND_EXC | O.COND_INH
    assert m.ampa_decay == 1.0 and m.gaba_decay == 1.0
    x = np.array([-55.0, ge, gi, I], np.float64)
    got = _run(m, x, 150, np.float64)
    v_inf = (r * I + ge * Ee + gi * Ei + v_rest) / (1.0 + ge + gi)
    h = (1.0 + ge + gi) * dt / tau
    fac = 1.0 - h if integ == O.EULER else _rk4_factor(h)
    np.testing.assert_allclose(got, _closed(-55.0, v_inf, fac, 150), rtol=1e-12)


@pytest.mark.parametrize("integ", [O.EULER, O.RK4])
def test_lif_exp_triple_conductance_closed_form(integ):
    # TimeDrivenLIF_Exponential_triple.cpp:79-105: tau dV/dt = R (ga (Ea - V) + gg (Eg - V) + I) + V_rest - V
    # without NMDA; time constants of 1e30 ms give decay factors of exactly 1.
    dt, tau, r, Ea, Eg, v_rest, I, ga, gg = 0.125, 10.0, 0.5, 0.0, -70.0, -60.0, 6.0, 0.8, 0.3
    m = model(model=O.LIF_EXP3, integrator=integ, dt=dt, tau=tau, r=r, e_ampa=Ea, e_gaba=Eg, v_rest=v_rest,
              v_th=100.0, ampa_tau=1e30, gaba_tau=1e30, nmda_tau=1e30)
    m.cond = ALL_COND
    L_.orc_model_prepare(C.byref(m))
    x = np.array([-75.0, ga, gg, 0.0, I], np.float64)
    got = _run(m, x, 150, np.float64)
    v_inf = (r * (ga * Ea + gg * Eg + I) + v_rest) / (1.0 + r * (ga + gg))
    h = (1.0 + r * (ga + gg)) * dt / tau
    fac = 1.0 - h if integ == O.EULER else _rk4_factor(h)
    np.testing.assert_allclose(got, _closed(-75.0, v_inf, fac, 150), rtol=1e-12)


def test_lif_exp_triple_nmda_current_vanishes_at_reversal():
    # the NMDA current g_nmda B(V) (E_ampa - V) is zero at V = E_ampa whatever B(V) is, and of the
    # same sign as E_ampa - V elsewhere (depolarising below the reversal potential)
    Ea = -10.0
    base = dict(model=O.LIF_EXP3, integrator=O.EULER, dt=0.1, tau=10.0, r=1.0, e_ampa=Ea, e_gaba=-70.0,
                v_rest=-60.0, v_th=100.0, ampa_tau=1e30, gaba_tau=1e30, nmda_tau=1e30)
    m = model(**base)
    m.cond = ALL_COND
    L_.orc_model_prepare(C.byref(m))

    def one(v, gn):
        x = np.array([v, 0.0, 0.0, gn, 0.0], np.float64)
        step(m, x, np.array([10000], np.int32))
        return x[0]
    assert one(Ea, 0.7) == one(Ea, 0.0)
    assert one(-40.0, 0.7) > one(-40.0, 0.0)
    assert one(20.0, 0.7) < one(20.0, 0.0)


def _izh(**kw):
    return model(model=O.IZH, integrator=O.EULER, dt=0.5, v_rest=-65.0, izh_r=1.0, e_ampa=0.0, e_gaba=-80.0,
                 ampa_tau=0.0, gaba_tau=0.0, izh_v_th=1000.0, a=[0.02], b=[0.2], c=[-65.0], d=[8.0], **kw)


def _izh_step(m, v, u, ga, gg):
    x = np.array([v, u, ga, gg, 0.0, 0.0], np.float64)
    step(m, x, np.array([10000], np.int32))
    return x


@pytest.mark.parametrize("integ", [O.EULER, O.RK4])
def test_izhikevich_conductance_currents_at_reversal(integ):
    # TimeDrivenIzhikevic_Exponential_Decay.cpp:72-97: I_tot = g_ampa (E_ampa - v) + g_gaba (E_gaba - v) + I.
    # At v = E the channel's current is zero (an Euler step from E does not see the channel; the
    # later RK4 stages leave E); below E it depolarises, above it hyperpolarises.
    m = _izh()
    m.integrator = integ
    E_a, E_g = 0.0, -80.0
    if integ == O.EULER:
        assert _izh_step(m, E_a, -13.0, 0.6, 0.0)[0] == _izh_step(m, E_a, -13.0, 0.0, 0.0)[0]
        assert _izh_step(m, E_g, -13.0, 0.0, 0.9)[0] == _izh_step(m, E_g, -13.0, 0.0, 0.0)[0]
    assert _izh_step(m, -65.0, -13.0, 0.6, 0.0)[0] > _izh_step(m, -65.0, -13.0, 0.0, 0.0)[0]  # AMPA excites
    assert _izh_step(m, -65.0, -13.0, 0.0, 0.9)[0] < _izh_step(m, -65.0, -13.0, 0.0, 0.0)[0]  # GABA inhibits
    assert _izh_step(m, -90.0, -13.0, 0.0, 0.9)[0] > _izh_step(m, -90.0, -13.0, 0.0, 0.0)[0]  # below E_gaba


def test_izhikevich_conductance_term_is_linear_and_scaled_by_r():
    # Euler: v' - v'_0 = dt * R (g_ampa (E_ampa - v) + g_gaba (E_gaba - v)) exactly (up to FP64 rounding);
    # u uses the old v, so it is unchanged by the conductances
    for R in (1.0, 2.5):
        m = _izh()
        m.izh_r = R
        v, u = -61.0, -12.0
        for ga, gg in ((0.3, 0.0), (0.0, 0.7), (0.3, 0.7), (1.9, 0.05)):
            x1, x0 = _izh_step(m, v, u, ga, gg), _izh_step(m, v, u, 0.0, 0.0)
            expect = 0.5 * R * (ga * (0.0 - v) + gg * (-80.0 - v))
            assert x1[0] - x0[0] == pytest.approx(expect, rel=1e-12)
            assert x1[1] == x0[1]


def test_conductance_decay_factors():
    # exp(-dt / tau) per step after the update (TimeDrivenLIF_Exponential_Decay.cpp:37-39,78-85;
    # TimeDrivenIzhikevic_Exponential_Decay.cpp:32-39): after n steps g = g0 exp(-n dt / tau)
    n = 40
    m = model(model=O.LIF_EXP, dt=0.1, tau=20.0, g_tau=5.0, v_rest=-65.0, v_th=100.0)
    x = np.array([-65.0, 1.0, 0.0], np.float32)
    last = np.array([10000], np.int32)
    for _ in range(n):
        step(m, x, last)
    assert float(x[1]) == pytest.approx(math.exp(-n * 0.1 / 5.0), rel=1e-5)
    mi = _izh()
    mi.ampa_tau, mi.gaba_tau = 4.0, 9.0
    L_.orc_model_prepare(C.byref(mi))
    xi = np.array([-65.0, -13.0, 1.0, 2.0, 0.0, 0.0], np.float32)
    last = np.array([10000], np.int32)
    for _ in range(n):
        step(mi, xi, last)
    assert float(xi[2]) == pytest.approx(math.exp(-n * 0.5 / 4.0), rel=1e-5)
    assert float(xi[3]) == pytest.approx(2.0 * math.exp(-n * 0.5 / 9.0), rel=1e-5)
