"""ctypes loader for the CPU oracle (test infrastructure only).

The oracle is a plain-C restatement of the reference's legacy CPU path
(see ``nbx_oracle.h`` for the reference file:line list). Only ``tests/``,
``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` import this
module, and only to check or time against the HIP library; the product package
``neuronbridge_amd`` never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")

LIF_VJ, LIF_EXP, IZH, IZH_CUR, LIF_EXP2, LIF_EXP3, POISSON, TRIGGER_RELAY = 0, 1, 2, 3, 4, 5, 6, 7
COND_EXC, COND_INH, COND_NMDA = 1, 2, 4
EULER, RK4 = 0, 1
ACC_FIXED, ACC_FLOAT_EVENTS = 0, 1

_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)
_u32p = C.POINTER(C.c_uint32)


class OrcModel(C.Structure):
    _fields_ = [
        ("model", C.c_int32), ("integrator", C.c_int32), ("n", C.c_int32), ("t_ref", C.c_int32),
        ("dt", C.c_float),
        ("v_rest", C.c_float), ("tau", C.c_float), ("v_th", C.c_float), ("r", C.c_float),
        ("v_reset", C.c_float), ("g_tau", C.c_float), ("e_rev", C.c_float),
        ("izh_v_th", C.c_float), ("izh_r", C.c_float), ("e_ampa", C.c_float), ("e_gaba", C.c_float),
        ("ampa_tau", C.c_float), ("gaba_tau", C.c_float),
        ("a", _f32p), ("b", _f32p), ("c", _f32p), ("d", _f32p),
        ("g_decay", C.c_float), ("ampa_decay", C.c_float), ("gaba_decay", C.c_float),
        ("nmda_tau", C.c_float), ("rate_bias", C.c_float), ("rate_gain", C.c_float),
        ("nmda_decay", C.c_float), ("cond", C.c_int32), ("step", C.c_int32),
    ]


class OrcNet(C.Structure):
    _fields_ = [
        ("m", OrcModel),
        ("d_slots", C.c_int32), ("fixed_shift", C.c_int32),
        ("row_ptr", _i64p), ("post", _i32p), ("delay", _u8p), ("type", _u8p), ("w", _f32p),
        ("dense_wt", _f32p), ("dense_delay", C.c_int32),
        ("pois_on", C.c_int32), ("seed_lo", C.c_uint32), ("seed_hi", C.c_uint32),
        ("pois_lambda", C.c_double), ("w_bg", C.c_float),
        ("cur", _f32p), ("cur_t_begin", C.c_int32),
        ("stdp_on", C.c_int32), ("plastic", _u8p),
        ("max_ltp", C.c_float), ("ltp_tau", C.c_float), ("max_ltd", C.c_float), ("ltd_tau", C.c_float),
        ("w_max", C.c_float),
        ("apre", _f32p), ("apost", _f32p), ("slast", _i32p),
        ("in_ptr", _i64p), ("in_syn", _i64p),
        ("nrules", C.c_int32), ("rules", C.c_void_p), ("syn_rule", C.POINTER(C.c_int8)),
        ("syn_trig", C.POINTER(C.c_int8)), ("maxw", _f32p), ("elig", _f32p), ("stdp_maxw", _f32p),
        ("conn_index", _i32p),
        ("npops", C.c_int32), ("pops", C.POINTER(OrcModel)), ("pop_begin", _i32p),
        ("inputs", C.c_void_p), ("input_ptr", _i32p), ("input_t0", C.c_int32), ("input_nsteps", C.c_int32),
    ]


class OrcInput(C.Structure):
    _fields_ = [("target", C.c_int32), ("kind", C.c_int32), ("value", C.c_float)]


INPUT_CURRENT = 3


RULE_RSTDP, RULE_ADDITIVE, RULE_CEREBELLAR = 1, 2, 3


class OrcRule(C.Structure):
    _fields_ = [("kind", C.c_int32), ("clear", C.c_int32), ("max_ltp", C.c_float), ("ltp_tau", C.c_float),
                ("max_ltd", C.c_float), ("ltd_tau", C.c_float), ("reward", C.c_float), ("punishment", C.c_float),
                ("a1pre", C.c_float), ("a2prepre", C.c_float), ("initpos", C.c_float), ("maxpos", C.c_float),
                ("kernel_step", C.c_float), ("min_weight", C.c_float), ("seed", C.c_uint32)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_nstate.restype = C.c_int
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_hash_uniform01.restype = C.c_float
        L.orc_hash_uniform01.argtypes = [C.c_int32, C.c_int32]
        L.orc_poisson_thresholds.restype = C.c_int32
        L.orc_poisson_thresholds.argtypes = [C.c_double, _u32p, C.c_int32]
        L.orc_poisson_count.restype = C.c_int32
        L.orc_poisson_count.argtypes = [C.c_uint32, C.c_uint32, C.c_int32, C.c_int32, _u32p, C.c_int32]
        L.orc_quantize.restype = C.c_int32
        L.orc_quantize.argtypes = [C.c_float, C.c_int32]
        L.orc_model_prepare.argtypes = [C.POINTER(OrcModel)]
        L.orc_init_state.argtypes = [C.POINTER(OrcModel), _f32p, _i32p]
        L.orc_update_f32.restype = C.c_int32
        L.orc_update_f32.argtypes = [C.POINTER(OrcModel), _f32p, _i32p, _f32p, _i32p]
        L.orc_update_f64.restype = C.c_int32
        L.orc_update_f64.argtypes = [C.POINTER(OrcModel), _f64p, _i32p, _f64p, _i32p]
        L.orc_run.restype = C.c_int32
        L.orc_run.argtypes = [C.POINTER(OrcNet), C.c_int32, _f32p, _i32p, _i32p, _i32p, _i32p,
                              C.c_int32, C.c_int32, _i32p, _u32p, _i64p]
        L.orc_stdp_replay.argtypes = [C.c_float] * 7 + [_u8p, _u8p, C.c_int32, _f32p]
        L.orc_replay_stdp.restype = C.c_int64
        L.orc_replay_stdp.argtypes = [C.POINTER(OrcNet), _i32p, _i32p, C.c_int32, C.c_int32, C.c_int32]
        L.orc_run_queues.restype = C.c_int32
        L.orc_run_queues.argtypes = [C.POINTER(OrcNet), C.c_int32, _f32p, _i32p, _i32p, _i32p, C.c_int32, C.c_int32,
                                     _i32p, _i64p, C.POINTER(C.c_double)]
        L.orc_stable_unit_random.restype = C.c_float
        L.orc_stable_unit_random.argtypes = [C.c_uint32, C.c_int32, C.c_int32, C.c_int32]
        L.orc_cerebellar_kernel.restype = C.c_int32
        L.orc_cerebellar_kernel.argtypes = [C.POINTER(OrcRule), _f32p, C.c_int32, _f32p, _f32p]
        _lib = L
    return _lib


def ptr(a: np.ndarray | None, ctype):
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data_as(C.POINTER(ctype))


def philox(ctr, key) -> np.ndarray:
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(ptr(c, C.c_uint32), ptr(k, C.c_uint32), ptr(out, C.c_uint32))
    return out


def nstate(model: int) -> int:
    return int(lib().orc_nstate(model))
