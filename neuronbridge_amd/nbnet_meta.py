"""Neuron populations, initial state and learning rules from a .nbnet file's metadata JSON.

The reference's streaming build reads the metadata with NbnetReader::ParseMetadata
(src/native/shared/src/streaming_build/NbnetReader.cpp:301-436: layers with model, count,
update_timestep, monitored, output, communication_input and parameters; learning_rules with name
and parameters; parameter values plain or tagged {"__neuronbridge_parameter__": kind, "value": v})
and hands each layer to the legacy model's SetParameters, each rule to LearningRuleModelFactory
(src/native/legacy/source_file_realtime_v1_async/ModelFactory/LearningRuleModelFactory.cpp:15-45).
This module maps the same JSON onto this library's descriptions, with the reference's parameter
names, the order in which each SetParameters reads them (a later alias overrides an earlier key) and
the defaults of the model and rule headers:

  TimeDrivenLIF_Voltage_jump        NeuralModel/inc/TimeDriven/TimeDrivenLIF_Voltage_jump.h:22-48
  TimeDrivenLIF_Exponential_Decay   TimeDrivenLIF_Exponential_Decay.h:24-56
  TimeDrivenLIF_Exponential_double  TimeDrivenLIF_Exponential_double.h:16-38
  TimeDrivenLIF_Exponential_triple  TimeDrivenLIF_Exponential_triple.h:13-29 (+ the double's, tau 10:
                                    TimeDrivenLIF_Exponential_triple.cpp:32-48,223-243)
  TimeDrivenIzhikevic_Exponential_Decay  TimeDrivenIzhikevic_Exponential_Decay.h:28-76
  PoissonRate                       PoissonRate.h:12-13
  STDP / R_STDP / AdditiveKernalChange / CerebullarLearningRule   LearningRule/inc/Rule/*.h

Initial state follows each model's InitializeStates (e.g. TimeDrivenIzhikevic_Exponential_Decay.cpp:49:
{V_rest + mu0, b * V_rest + mu1, mu2, ...}, float arithmetic) with random_sigma all zero; a non-zero
sigma draws from the reference's own random generator, which this library does not reproduce, and is
refused. Every layer steps with the simulation's dt (update_timestep must be 1); the legacy models
integrate with ForwardEulerMethod unless the caller asks for RK4.
"""
from __future__ import annotations

import json
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .network import ModelSpec, PopulationsSpec

MARKER = "__neuronbridge_parameter__"
f32 = np.float32


def parse_parameters(params: dict | None) -> dict:
    """Parameter values as NbnetReader reads them: tagged values by their kind, plain values as
    bool / int / float / string (ParseLegacyParameter), plain lists as lists."""
    out = {}
    for key, v in (params or {}).items():
        if isinstance(v, dict) and MARKER in v:
            kind, val = v[MARKER], v["value"]
            if kind == "bool":
                out[key] = bool(val)
            elif kind in ("int", "int32"):
                out[key] = int(val)
            elif kind in ("float", "float32"):
                out[key] = f32(val)
            elif kind == "float64":
                out[key] = float(val)
            elif kind == "str":
                out[key] = str(val)
            elif kind.endswith("_list") or kind.startswith("float32_array"):
                if kind.startswith("float32_array") and len(val) != int(kind[-1]):
                    raise ValueError(f"{kind} requires {kind[-1]} values")
                out[key] = list(val)
            else:
                raise ValueError(f"unsupported nbnet parameter kind: {kind}")
        else:
            out[key] = v
    return out


# model name -> (nbx model, the state columns' random_mu length, parameter table)
# parameter table: field -> (default, [keys in the order SetParameters reads them], type)
_LIF_COMMON = {"r": (1.0, ["R"], f32), "t_ref": (50, ["t_ref"], int)}
MODELS = {
    "TimeDrivenLIF_Voltage_jump": (L.LIF_VJ, 2, dict(
        v_rest=(0.0, ["V_rest"], f32), tau=(10.0, ["tau"], f32), v_th=(20.0, ["V_th"], f32),
        v_reset=(-5.0, ["V_reset"], f32), **_LIF_COMMON)),
    "TimeDrivenLIF_Exponential_Decay": (L.LIF_EXP, 3, dict(
        v_rest=(-65.0, ["V_rest"], f32), tau=(20.0, ["tau"], f32), v_th=(-20.0, ["V_th"], f32),
        v_reset=(-65.0, ["V_reset"], f32), g_tau=(12.0, ["g_tau"], f32), e_rev=(0.0, ["E"], f32),
        **_LIF_COMMON)),
    "TimeDrivenLIF_Exponential_double": (L.LIF_EXP2, 4, dict(
        v_rest=(-60.0, ["V_rest"], f32), tau=(20.0, ["tau"], f32), v_th=(-50.0, ["V_th"], f32),
        v_reset=(-60.0, ["V_reset"], f32), ampa_tau=(5.0, ["gexc_tau"], f32), e_ampa=(0.0, ["Eexc"], f32),
        gaba_tau=(10.0, ["ginh_tau"], f32), e_gaba=(-80.0, ["Einh"], f32), **_LIF_COMMON)),
    "TimeDrivenLIF_Exponential_triple": (L.LIF_EXP3, 5, dict(
        v_rest=(-60.0, ["V_rest"], f32), tau=(10.0, ["tau"], f32), v_th=(-50.0, ["V_th"], f32),
        v_reset=(-60.0, ["V_reset"], f32), e_ampa=(0.0, ["E_ampa", "Eexc"], f32),
        ampa_tau=(10.0, ["ampa_tau", "gexc_tau"], f32), e_gaba=(-80.0, ["E_gaba", "Einh"], f32),
        gaba_tau=(10.0, ["gaba_tau", "ginh_tau"], f32), nmda_tau=(10.0, ["nmda_tau"], f32), **_LIF_COMMON)),
    "TimeDrivenIzhikevic_Exponential_Decay": (L.IZH, 6, dict(
        a=(0.02, ["a"], f32), b=(0.2, ["b"], f32), c=(-65.0, ["c"], f32), d=(8.0, ["d"], f32),
        v_rest=(-65.0, ["V_rest"], f32), v_reset=(-65.0, ["V_reset"], f32), izh_v_th=(30.0, ["V_th"], f32),
        izh_r=(1.0, ["R"], f32), e_ampa=(0.0, ["E_ampa"], f32), ampa_tau=(5.0, ["ampa_tau"], f32),
        e_gaba=(-80.0, ["E_gaba"], f32), gaba_tau=(10.0, ["gaba_tau"], f32), nmda_tau=(100.0, ["nmda_tau"], f32))),
    "PoissonRate": (L.POISSON, 2, dict(
        rate_bias_hz=(0.0, ["poisson_rate_bias_hz", "rate_bias_hz"], f32),
        rate_gain_hz=(1.0, ["poisson_rate_gain_hz_per_current", "rate_gain_hz_per_current"], f32))),
}
# keys every layer may carry that do not change the neuron update here
_IGNORED = {"dense_subnetwork_name", "int_method", "random_mu", "random_sigma"}


def _read(params: dict, table: dict) -> dict:
    vals = {}
    for fld, (default, keys, typ) in table.items():
        v = default
        for k in keys:  # the reference reads the keys in this order: a later one overrides
            if k in params:
                v = params[k]
        vals[fld] = typ(v)
    unknown = set(params) - {k for _, keys, _ in table.values() for k in keys} - _IGNORED
    if unknown:
        raise ValueError(f"parameters this model does not take: {sorted(unknown)}")
    return vals


def _initial_state(model: int, vals: dict, mu: list, n: int) -> np.ndarray:
    """InitializeStates of each model with sigma 0 (float arithmetic, the reference's order)."""
    mu = [f32(x) for x in mu]
    first = {L.LIF_VJ: vals.get("v_reset"), L.LIF_EXP: vals.get("v_rest"), L.LIF_EXP2: vals.get("v_rest"),
             L.LIF_EXP3: vals.get("v_reset"), L.IZH: vals.get("v_rest")}.get(model)
    row = list(mu)
    if first is not None:
        row[0] = f32(f32(first) + mu[0])
    if model == L.IZH:
        row[1] = f32(f32(vals["b"] * vals["v_rest"]) + mu[1])
    return np.tile(np.asarray(row, np.float32), (n, 1))


@dataclass
class Layer:
    name: str
    model: int
    begin: int
    count: int
    monitored: bool
    output: bool
    communication_input: bool
    dense_subnetwork_name: str = ""


@dataclass
class NbnetDescription:
    """What the metadata describes: the populations (ModelSpec for one layer, PopulationsSpec for
    several), their initial state, the output mask (layers marked output), the monitored neurons,
    the learning rules (dicts for Net.assign_nbnet_rules, by the records' rule ids) and the layers."""
    spec: object
    x0: np.ndarray
    last0: np.ndarray
    output_mask: np.ndarray | None
    monitored: np.ndarray
    rules: list
    layers: list = field(default_factory=list)


def describe(meta, dt: float, integrator: int = L.EULER) -> NbnetDescription:
    """Maps metadata JSON (bytes, str or an already parsed dict) onto populations and rules."""
    if isinstance(meta, (bytes, bytearray)):
        meta = meta.decode()
    if isinstance(meta, str):
        meta = json.loads(meta)
    layers_in = meta.get("layers")
    if not layers_in:
        raise ValueError("nbnet metadata has no layers")
    pops, states, layers = [], [], []
    begin = 0
    for k, ly in enumerate(layers_in):
        name = ly["model"]
        count = int(ly["count"])
        step = int(ly.get("update_timestep", 1))
        if count <= 0 or step <= 0:
            raise ValueError("invalid nbnet layer count or timestep")  # NbnetReader.cpp:419-421
        if step != 1:
            raise ValueError(f"layer {k}: update_timestep {step} (every population steps with the same dt)")
        if name not in MODELS:
            raise ValueError(f"layer {k}: model {name!r} is not supported")
        model, nst, table = MODELS[name]
        params = parse_parameters(ly.get("parameters"))
        vals = _read(params, table)
        sigma = params.get("random_sigma")
        if sigma is not None and any(float(s) != 0.0 for s in sigma):
            raise ValueError(f"layer {k}: random_sigma draws from the reference's generator; give the state instead")
        mu = params.get("random_mu", [0.0] * nst)
        if len(mu) != nst:
            raise ValueError(f"layer {k}: random_mu needs {nst} values for {name}")
        spec = ModelSpec(model=model, n=count, dt=float(dt), integrator=integrator)
        for fld, v in vals.items():
            if model == L.IZH and fld in "abcd":
                setattr(spec, fld, np.full(count, v, np.float32))
            elif fld == "t_ref":
                spec.t_ref = int(v)
            else:
                setattr(spec, fld, float(v))
        pops.append(spec)
        states.append(_initial_state(model, vals, mu, count))
        layers.append(Layer(name=name, model=model, begin=begin, count=count, monitored=bool(ly.get("monitored", False)),
                            output=bool(ly.get("output", False)),
                            communication_input=bool(ly.get("communication_input", False)),
                            dense_subnetwork_name=str(params.get("dense_subnetwork_name", ""))))
        begin += count
    n = begin
    spec = pops[0] if len(pops) == 1 else PopulationsSpec(pops)
    ns = spec.nstate
    x0 = np.zeros((n, ns), np.float32)
    for ly, st in zip(layers, states):
        x0[ly.begin:ly.begin + ly.count, :st.shape[1]] = st
    mask = None
    if any(ly.output for ly in layers):
        mask = np.zeros(n, np.uint8)
        for ly in layers:
            if ly.output:
                mask[ly.begin:ly.begin + ly.count] = 1
    monitored = np.concatenate([np.arange(ly.begin, ly.begin + ly.count, dtype=np.int32)
                                for ly in layers if ly.monitored] or [np.zeros(0, np.int32)])
    rules = [rule_desc(r.get("name", ""), parse_parameters(r.get("parameters"))) for r in meta.get("learning_rules", [])]
    return NbnetDescription(spec=spec, x0=x0, last0=np.full(n, 10000, np.int32), output_mask=mask,
                            monitored=monitored, rules=rules, layers=layers)


def rule_desc(name: str, p: dict) -> dict:
    """One learning rule as a dict for Net.assign_nbnet_rules (reference names, keys and defaults)."""
    def get(keys, default, typ=f32):
        v = default
        for k in keys:
            if k in p:
                v = p[k]
        return typ(v)
    stdp = dict(max_ltp=get(["Max_LTP"], 0.92), ltp_tau=get(["LTP_tau"], 16.8), max_ltd=get(["Max_LTD"], 0.53),
                ltd_tau=get(["LTD_tau"], 33.1))
    if name == "STDP":  # STDP.h:14-20
        return dict(kind=L.RULE_STDP, **{k: float(v) for k, v in stdp.items()})
    if name in ("R_STDP", "R-STDP"):  # R_STDP.h:8-20
        return dict(kind=L.RULE_RSTDP, **{k: float(v) for k, v in stdp.items()},
                    reward=float(get(["RewardFactor"], 1.0)), punishment=float(get(["PunishmentFactor"], -1.0)),
                    clear_eligibility=int(bool(get(["ClearEligibilityAfterTrigger"], True, bool))))
    # AdditiveKernalChange.h:8-11 and AdditiveKernalChange.cpp:23-44 (fixweightchange / kernalchange
    # override a1pre / a2prepre)
    additive = dict(a1pre=float(get(["a1pre", "fixweightchange"], 0.002)),
                    a2prepre=float(get(["a2prepre", "kernalchange"], -0.08)), ltp_tau=float(get(["LTP_tau"], 16.8)))
    if name == "AdditiveKernalChange":
        return dict(kind=L.RULE_ADDITIVE, **additive)
    if name == "CerebullarLearningRule":  # CerebullarLearningRule.h:9-21, .cpp:136-166
        return dict(kind=L.RULE_CEREBELLAR, **additive, initpos=float(get(["initpos"], 120.0)),
                    maxpos=float(get(["maxpos"], 150.0)), kernel_step=float(get(["kernel_step_size"], 0.1)),
                    min_weight=float(get(["min_weight", "minimum_weight", "w_gcpc_min"], 0.0)),
                    seed=int(get(["random_seed"], 17, int)) & 0xFFFFFFFF)
    raise ValueError(f"learning rule {name!r} is not supported")


def load_network(path: str, dt: float, integrator: int = L.EULER, device: int = 0, verify_checksum: bool = True,
                 poisson: tuple | None = None):
    """A finalized Net from a .nbnet file alone: populations and initial state from its metadata,
    the connection table loaded on the device, its learning rules assigned by the records' rule ids,
    the output mask of the layers marked output. poisson = (seed, lambda, w_bg) adds the background
    drive of nbx_net_set_poisson. Returns (net, description)."""
    from .network import Net, nbnet_metadata
    desc = describe(nbnet_metadata(path), dt, integrator)
    net = Net(desc.spec, device)
    net.load_nbnet(path, verify_checksum)
    if poisson is not None:
        net.set_poisson(*poisson)
    if desc.rules:
        net.assign_nbnet_rules(desc.rules)
    net.finalize()
    net.set_state(desc.x0, desc.last0)
    if desc.output_mask is not None:
        net.set_output_mask(desc.output_mask)
    return net, desc
