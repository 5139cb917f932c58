"""Metadata of .nbnet files onto populations and rules: the reference's parameter names, the order
its SetParameters reads them, header defaults, tagged values, initial states and rejections (CPU).
Defaults and keys are those of the legacy headers named in neuronbridge_amd/nbnet_meta.py."""
import json

import numpy as np
import pytest

from neuronbridge_amd import _lib as L
from neuronbridge_amd import nbnet_meta as M
from neuronbridge_amd.network import ModelSpec, PopulationsSpec


def tag(kind, value):
    return {M.MARKER: kind, "value": value}


def test_defaults_of_every_model():
    want = {
        "TimeDrivenLIF_Voltage_jump": dict(v_rest=0.0, tau=10.0, v_th=20.0, r=1.0, v_reset=-5.0, t_ref=50),
        "TimeDrivenLIF_Exponential_Decay": dict(v_rest=-65.0, tau=20.0, v_th=-20.0, v_reset=-65.0, g_tau=12.0,
                                                e_rev=0.0, t_ref=50),
        "TimeDrivenLIF_Exponential_double": dict(v_rest=-60.0, tau=20.0, v_th=-50.0, v_reset=-60.0, ampa_tau=5.0,
                                                 e_ampa=0.0, gaba_tau=10.0, e_gaba=-80.0, t_ref=50),
        "TimeDrivenLIF_Exponential_triple": dict(v_rest=-60.0, tau=10.0, v_th=-50.0, v_reset=-60.0, e_ampa=0.0,
                                                 ampa_tau=10.0, e_gaba=-80.0, gaba_tau=10.0, nmda_tau=10.0),
        "TimeDrivenIzhikevic_Exponential_Decay": dict(v_rest=-65.0, izh_v_th=30.0, izh_r=1.0, e_ampa=0.0,
                                                      ampa_tau=5.0, e_gaba=-80.0, gaba_tau=10.0, nmda_tau=100.0),
        "PoissonRate": dict(rate_bias_hz=0.0, rate_gain_hz=1.0),
    }
    for name, fields in want.items():
        d = M.describe({"layers": [{"model": name, "count": 3}]}, dt=0.1)
        assert isinstance(d.spec, ModelSpec) and d.spec.model == M.MODELS[name][0]
        for f, v in fields.items():
            assert getattr(d.spec, f) == pytest.approx(v), (name, f)
    izh = M.describe({"layers": [{"model": "TimeDrivenIzhikevic_Exponential_Decay", "count": 4}]}, dt=0.5).spec
    assert np.all(izh.a == np.float32(0.02)) and np.all(izh.d == np.float32(8.0)) and izh.a.size == 4


def test_parameters_plain_tagged_and_aliases():
    layer = {"model": "TimeDrivenLIF_Exponential_triple", "count": 2, "parameters": {
        "V_rest": tag("float32", -70.0), "t_ref": tag("int", 20), "E_ampa": 5.0, "Eexc": tag("float32", -3.0),
        "gaba_tau": 7.0, "ginh_tau": 9.0, "nmda_tau": tag("float32", 80.0), "dense_subnetwork_name": "d0"}}
    s = M.describe({"layers": [layer]}, dt=0.1).spec
    # the reference reads E_ampa then Eexc, gaba_tau then ginh_tau: the alias read later wins
    assert (s.v_rest, s.t_ref, s.e_ampa, s.gaba_tau, s.nmda_tau) == (-70.0, 20, -3.0, 9.0, 80.0)
    p = M.describe({"layers": [{"model": "PoissonRate", "count": 1, "parameters": {
        "rate_bias_hz": 5.0, "poisson_rate_bias_hz": 7.0, "rate_gain_hz_per_current": 3.0}}]}, dt=0.1).spec
    assert (p.rate_bias_hz, p.rate_gain_hz) == (5.0, 3.0)  # poisson_rate_bias_hz is read first


def test_initial_state_formulas():
    mu = [1.5, 0.25, 0.0, 0.0, 0.0, 2.0]
    d = M.describe({"layers": [
        {"model": "TimeDrivenLIF_Voltage_jump", "count": 2, "parameters": {"V_reset": -7.0, "random_mu": [3.0, 0.5]}},
        {"model": "TimeDrivenIzhikevic_Exponential_Decay", "count": 3,
         "parameters": {"V_rest": -61.0, "b": 0.25, "random_mu": tag("float32_list", mu),
                        "random_sigma": [0, 0, 0, 0, 0, 0]}},
        {"model": "TimeDrivenLIF_Exponential_Decay", "count": 1, "parameters": {"V_rest": -66.0}},
    ]}, dt=0.1)
    assert isinstance(d.spec, PopulationsSpec) and d.x0.shape == (6, 6)
    assert list(d.x0[0, :2]) == [np.float32(-4.0), np.float32(0.5)]          # V_reset + mu0, mu1
    u = np.float32(np.float32(np.float32(0.25) * np.float32(-61.0)) + np.float32(0.25))
    assert d.x0[2, 0] == np.float32(-59.5) and d.x0[2, 1] == u and d.x0[2, 5] == 2.0  # V_rest + mu0, b V_rest + mu1
    assert d.x0[5, 0] == np.float32(-66.0)                                     # V_rest + 0
    assert np.all(d.last0 == 10000)
    assert [(ly.begin, ly.count) for ly in d.layers] == [(0, 2), (2, 3), (5, 1)]


def test_output_mask_monitored_and_rules():
    meta = {"layers": [{"model": "TimeDrivenLIF_Voltage_jump", "count": 3, "output": True},
                       {"model": "TimeDrivenLIF_Voltage_jump", "count": 2, "monitored": True}],
            "learning_rules": [
                {"name": "STDP", "parameters": {"Max_LTP": 0.01, "LTP_tau": 20.0}},
                {"name": "R-STDP", "parameters": {"RewardFactor": 0.5, "ClearEligibilityAfterTrigger": False}},
                {"name": "AdditiveKernalChange", "parameters": {"a1pre": 1e-3, "fixweightchange": 2e-3}},
                {"name": "CerebullarLearningRule", "parameters": {"initpos": 2.0, "maxpos": 6.0,
                                                                  "min_weight": 0.1, "w_gcpc_min": 0.2,
                                                                  "random_seed": tag("int", 5)}}]}
    d = M.describe(json.dumps(meta).encode(), dt=1.0)
    assert list(d.output_mask) == [1, 1, 1, 0, 0] and list(d.monitored) == [3, 4]
    stdp, rs, add, cb = d.rules
    assert stdp["kind"] == L.RULE_STDP and stdp["max_ltp"] == pytest.approx(0.01) and stdp["ltd_tau"] == pytest.approx(33.1)
    assert rs["kind"] == L.RULE_RSTDP and rs["reward"] == 0.5 and rs["punishment"] == -1.0 and rs["clear_eligibility"] == 0
    assert rs["max_ltp"] == pytest.approx(0.92)
    assert add["kind"] == L.RULE_ADDITIVE and add["a1pre"] == pytest.approx(2e-3) and add["a2prepre"] == pytest.approx(-0.08)
    assert cb["kind"] == L.RULE_CEREBELLAR and (cb["initpos"], cb["maxpos"], cb["seed"]) == (2.0, 6.0, 5)
    assert cb["min_weight"] == pytest.approx(0.2) and cb["kernel_step"] == pytest.approx(0.1)
    assert cb["a1pre"] == pytest.approx(0.002)


@pytest.mark.parametrize("layer,msg", [
    ({"model": "TimeDrivenLIF_Voltage_jump", "count": 0}, "count or timestep"),
    ({"model": "TimeDrivenLIF_Voltage_jump", "count": 2, "update_timestep": 2}, "update_timestep"),
    ({"model": "NoSuchModel", "count": 2}, "not supported"),
    ({"model": "TimeDrivenLIF_Voltage_jump", "count": 2, "parameters": {"random_sigma": [0.0, 1.0]}}, "random_sigma"),
    ({"model": "TimeDrivenLIF_Voltage_jump", "count": 2, "parameters": {"g_tau": 3.0}}, "does not take"),
    ({"model": "TimeDrivenLIF_Voltage_jump", "count": 2, "parameters": {"random_mu": [1.0]}}, "random_mu"),
    ({"model": "TimeDrivenLIF_Voltage_jump", "count": 2, "parameters": {"V_th": tag("complex", 1)}}, "kind"),
])
def test_rejections(layer, msg):
    with pytest.raises(ValueError, match=msg):
        M.describe({"layers": [layer]}, dt=0.1)
    with pytest.raises(ValueError, match="not supported"):
        M.rule_desc("NoSuchRule", {})
