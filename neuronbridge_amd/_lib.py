"""ctypes binding of libnbx.so (the C ABI in include/nbx_core.h).

Loading fails loudly when the HIP library has not been built: there is no
CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libnbx.so")

NBX_OK = 0
NBX_ERR_INVALID = -1
NBX_ERR_HIP = -2
NBX_ERR_UNSUPPORTED = -3
NBX_ERR_STATE = -4

LIF_VJ, LIF_EXP, IZH, IZH_CUR, LIF_EXP2, LIF_EXP3, POISSON, TRIGGER_RELAY = 0, 1, 2, 3, 4, 5, 6, 7
EULER, RK4 = 0, 1
ENGINE_KERNELS, ENGINE_GRAPH, ENGINE_PERSISTENT, ENGINE_AUTO = 0, 1, 2, 3
REPLAY_GLOBAL, REPLAY_OFFSETS, REPLAY_SLOTS, REPLAY_STREAM = 0, 1, 2, 3
RECORD_COUNTS, RECORD_RASTER = 1, 2

_f32p = C.POINTER(C.c_float)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)
_u32p = C.POINTER(C.c_uint32)
_vp = C.c_void_p


class NbxError(RuntimeError):
    """A libnbx call returned a non-zero status (message from nbx_last_error)."""

    def __init__(self, fn: str, status: int, message: str):
        super().__init__(f"{fn} failed with status {status}: {message}")
        self.status = status


class ModelDesc(C.Structure):
    _fields_ = [
        ("model", C.c_int32), ("integrator", C.c_int32), ("n", C.c_int32), ("t_ref", C.c_int32),
        ("dt", C.c_float),
        ("v_rest", C.c_float), ("tau", C.c_float), ("v_th", C.c_float), ("r", C.c_float),
        ("v_reset", C.c_float), ("g_tau", C.c_float), ("e_rev", C.c_float),
        ("izh_v_th", C.c_float), ("izh_r", C.c_float), ("e_ampa", C.c_float), ("e_gaba", C.c_float),
        ("ampa_tau", C.c_float), ("gaba_tau", C.c_float),
        ("izh_a", _f32p), ("izh_b", _f32p), ("izh_c", _f32p), ("izh_d", _f32p),
        ("nmda_tau", C.c_float), ("rate_bias_hz", C.c_float), ("rate_gain_hz", C.c_float),
    ]


RULE_STDP, RULE_RSTDP, RULE_ADDITIVE, RULE_CEREBELLAR = 0, 1, 2, 3


class RuleDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("max_ltp", C.c_float), ("ltp_tau", C.c_float), ("max_ltd", C.c_float),
                ("ltd_tau", C.c_float), ("reward", C.c_float), ("punishment", C.c_float),
                ("clear_eligibility", C.c_int32), ("a1pre", C.c_float), ("a2prepre", C.c_float),
                ("initpos", C.c_float), ("maxpos", C.c_float), ("kernel_step", C.c_float), ("min_weight", C.c_float),
                ("seed", C.c_uint32), ("history", C.c_int32)]


def model_channels(model: int) -> int:
    """Arrival channels of a model's delay ring (libnbx's model_channels)."""
    return 3 if model == LIF_EXP3 else (2 if model in (IZH, LIF_EXP2, TRIGGER_RELAY) else 1)


def model_nstate(model: int) -> int:
    """Columns of the reference's AoS state per model."""
    return {LIF_VJ: 2, LIF_EXP: 3, LIF_EXP2: 4, LIF_EXP3: 5, POISSON: 2, TRIGGER_RELAY: 1}.get(model, 6)


class SparseDesc(C.Structure):
    _fields_ = [
        ("count", C.c_int64), ("max_delay", C.c_int32),
        ("row_ptr", _i64p), ("post", _i32p), ("delay", _u8p), ("type", _u8p), ("weight", _f32p),
        ("plastic", _u8p),
    ]


class StdpDesc(C.Structure):
    _fields_ = [("max_ltp", C.c_float), ("ltp_tau", C.c_float), ("max_ltd", C.c_float),
                ("ltd_tau", C.c_float), ("w_max", C.c_float)]


class BatchInstance(C.Structure):
    _fields_ = [("n", C.c_int32), ("tau", C.c_float), ("pois_lambda", C.c_double), ("seed_lo", C.c_uint32),
                ("seed_hi", C.c_uint32), ("fixed_shift", C.c_int32), ("w_bf16", C.POINTER(C.c_uint16)),
                ("v0", C.POINTER(C.c_float))]


class BatchDesc(C.Structure):
    _fields_ = [("count", C.c_int32), ("integrator", C.c_int32), ("dt", C.c_float), ("t_ref", C.c_int32),
                ("v_rest", C.c_float), ("v_th", C.c_float), ("r", C.c_float), ("v_reset", C.c_float),
                ("w_bg", C.c_float), ("inst", C.POINTER(BatchInstance))]


class LifAdjParams(C.Structure):
    _fields_ = [("n", C.c_int32), ("batch", C.c_int32), ("g", C.c_float), ("one_minus_g", C.c_float),
                ("r", C.c_float), ("v_rest", C.c_float), ("v_th", C.c_float), ("delta_reset", C.c_float),
                ("slope", C.c_float), ("mode", C.c_int32)]


class PendingView(C.Structure):
    _fields_ = [("d_pending", _vp), ("channel_count", C.c_int32), ("channel_stride", C.c_int32),
                ("neuron_count", C.c_int32), ("inv_scale", C.c_float)]


class NbnetHeader(C.Structure):
    _fields_ = [("file_bytes", C.c_int64), ("metadata_bytes", C.c_int64), ("connection_count", C.c_int64),
                ("version_major", C.c_int32), ("version_minor", C.c_int32), ("flags", C.c_int32),
                ("checksum_verified", C.c_int32)]


class MonitorDesc(C.Structure):
    _fields_ = [("output_dir", C.c_char_p), ("component_name", C.c_char_p),
                ("sample_interval_steps", C.c_int32), ("flush_interval_steps", C.c_int32),
                ("record_spikes", C.c_int32), ("record_state", C.c_int32), ("record_weights", C.c_int32),
                ("record_pending_channels", C.c_int32), ("neuron_ids", _i32p), ("neuron_count", C.c_int32)]


class InterfaceValue(C.Structure):
    _fields_ = [("interface_slot", C.c_int32), ("target_neuron", C.c_int32), ("pending_channel", C.c_uint8),
                ("overwrite", C.c_uint8), ("value", C.c_float), ("scale", C.c_float)]


class InterfaceCopy(C.Structure):
    _fields_ = [("interface_slot", C.c_int32), ("target_neuron", C.c_int32), ("pending_channel", C.c_uint8),
                ("overwrite", C.c_uint8), ("source_index", C.c_int32), ("scale", C.c_float)]


PENDING_EXC, PENDING_INH, PENDING_CURRENT, PENDING_NMDA, PENDING_MODULATORY = 0, 1, 2, 3, 4

_NET = _vp
# name -> argtypes (all return int32 status unless listed in _RESTYPES)
SIGNATURES = {
    "nbx_last_error": [],
    "nbx_version": [],
    "nbx_device_count": [],
    "nbx_net_create": [C.POINTER(ModelDesc), C.c_int32, C.POINTER(_vp)],
    "nbx_net_create_populations": [C.POINTER(ModelDesc), C.c_int32, C.c_int32, C.POINTER(_vp)],
    "nbx_net_set_output_mask": [_NET, _u8p],
    "nbx_net_neuron_step_outputs": [_NET, _i32p, C.c_int32, _i32p, _i32p, C.c_int32, _i32p],
    "nbx_net_set_interface_targets": [_NET, _i32p, C.c_int32],
    "nbx_net_upload_interface_values": [_NET, C.POINTER(InterfaceValue), C.c_int32],
    "nbx_net_apply_interface": [_NET],
    "nbx_net_upload_interface_values_device": [_NET, _vp, C.POINTER(InterfaceCopy), C.c_int32],
    "nbx_net_add_arrivals": [_NET, _i32p, _u8p, _f32p, C.c_int32, C.c_int32],
    "nbx_net_reset_state": [_NET],
    "nbx_net_download_pending": [_NET, _f32p, C.c_int64, _i32p, _i32p],
    "nbx_net_pending_device_float": [_NET, C.POINTER(_vp)],
    "nbx_net_commit_external": [_NET, C.c_int32, _vp, _vp],
    "nbx_net_firing_device_view": [_NET, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)],
    "nbx_net_destroy": [_NET],
    "nbx_net_set_sparse": [_NET, C.POINTER(SparseDesc)],
    "nbx_net_set_dense": [_NET, _f32p, C.c_int32],
    "nbx_net_set_poisson": [_NET, C.c_uint32, C.c_uint32, C.c_double, C.c_float],
    "nbx_net_set_stdp": [_NET, C.POINTER(StdpDesc)],
    "nbx_net_set_current": [_NET, _f32p, C.c_int32, C.c_int32],
    "nbx_net_set_fixed_shift": [_NET, C.c_int32],
    "nbx_net_get_fixed_shift": [_NET, _i32p],
    "nbx_net_finalize": [_NET],
    "nbx_net_nstate": [_NET],
    "nbx_net_set_state": [_NET, _f32p, _i32p],
    "nbx_net_get_state": [_NET, _f32p, _i32p],
    "nbx_net_record": [_NET, C.c_int32, C.c_int32, C.c_int32],
    "nbx_net_get_counts": [_NET, _i32p, C.c_int32],
    "nbx_net_get_raster": [_NET, _u32p, C.c_int64],
    "nbx_net_run": [_NET, C.c_int32, C.c_int32],
    "nbx_net_run_async": [_NET, C.c_int32, C.c_int32],
    "nbx_net_sync": [_NET],
    "nbx_net_pick_engine": [_NET, _i32p],
    "nbx_net_time": [_NET, _i32p],
    "nbx_net_events": [_NET, _i64p],
    "nbx_net_begin_step": [_NET, C.c_int32],
    "nbx_net_run_propagation": [_NET],
    "nbx_net_neuron_step": [_NET, _i32p, C.c_int32, _i32p],
    "nbx_net_apply_post_learning": [_NET],
    "nbx_net_commit_firings": [_NET],
    "nbx_net_inject_spikes": [_NET, _i32p, C.c_int32],
    "nbx_net_ring_shape": [_NET, _i32p, _i32p],
    "nbx_net_get_ring": [_NET, _i32p, C.c_int64],
    "nbx_net_get_history": [_NET, C.c_int32, _i32p, C.c_int32, _i32p],
    "nbx_net_pending_view": [_NET, C.POINTER(PendingView)],
    "nbx_net_get_weights": [_NET, _f32p, C.c_int64],
    "nbx_net_set_weights": [_NET, _f32p, C.c_int64],
    "nbx_net_get_synapse_weight": [_NET, C.c_int64, _f32p],
    "nbx_net_set_synapse_weight": [_NET, C.c_int64, C.c_float],
    "nbx_net_get_stdp_state": [_NET, _f32p, _f32p, _i32p, C.c_int64],
    "nbx_replay_deliver": [_NET, _i32p, _i32p, C.c_int32, C.c_int32, _i32p],
    "nbx_replay_deliver_device": [_NET, _vp, _vp, C.c_int32, _vp],
    "nbx_replay_path": [_NET, _i32p],
    "nbx_replay_stdp": [_NET, _i32p, _i32p, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_int64)],
    "nbx_replay_stdp_device": [_NET, _vp, _vp, C.c_int32, C.c_int32, C.c_int32, C.c_int64, _vp],
    "nbx_replay_stdp_device_noalloc": [_NET, _vp, _vp, C.c_int32, C.c_int32, C.c_int32, C.c_int64, _vp],
    "nbx_nbnet_info": [C.c_char_p, C.c_int32, C.POINTER(NbnetHeader)],
    "nbx_nbnet_metadata": [C.c_char_p, C.c_char_p, C.c_int64, C.POINTER(C.c_int64)],
    "nbx_net_load_nbnet": [_NET, C.c_char_p, C.c_int32],
    "nbx_net_export_monitor": [_NET, C.POINTER(MonitorDesc)],
    "nbx_net_set_current_synapses": [_NET, _vp, _vp, _vp, C.c_int32],
    "nbx_net_inject_currents": [_NET, _vp, _vp, C.c_int32, C.c_int32],
    "nbx_net_monitor_start": [_NET, C.POINTER(MonitorDesc)],
    "nbx_net_monitor_flush": [_NET],
    "nbx_net_monitor_stop": [_NET],
    "nbx_net_monitor_frames": [_NET, C.POINTER(C.c_int64), C.POINTER(C.c_int32)],
    "nbx_net_set_rules": [_NET, C.POINTER(RuleDesc), C.c_int32, _i32p, _i32p, _f32p],
    "nbx_net_get_rule_state": [_NET, _f32p, _f32p, _f32p, _i32p, C.c_int64],
    "nbx_net_assign_nbnet_rules": [_NET, C.POINTER(RuleDesc), C.c_int32],
    "nbx_build_fixed_indegree": [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint32, C.c_uint32,
                                 C.c_int32, C.c_int32, _i64p, _i32p, _u8p, _u8p],
    "nbx_net_stream": [_NET],
    "nbx_build_dense_normal_bf16": [C.c_int32, C.c_float, C.c_uint32, C.c_uint32, C.c_int32, C.POINTER(C.c_uint16)],
    "nbx_batch_create": [C.POINTER(BatchDesc), C.c_int32, C.POINTER(_vp)],
    "nbx_batch_destroy": [_vp],
    "nbx_batch_tune": [_vp, C.c_int32, _i32p],
    "nbx_batch_set_schedule": [_vp, C.c_int32, C.c_int32],
    "nbx_batch_get_schedule": [_vp, _i32p, _i32p],
    "nbx_batch_set_parts": [_vp, _i32p, C.c_int32],
    "nbx_batch_get_parts": [_vp, _i32p, C.c_int32],
    "nbx_batch_record": [_vp, C.c_int32, C.c_int32],
    "nbx_batch_run": [_vp, C.c_int32],
    "nbx_batch_run_async": [_vp, C.c_int32],
    "nbx_batch_sync": [_vp],
    "nbx_batch_get_raster": [_vp, C.c_int32, _u32p, C.c_int64],
    "nbx_batch_get_state": [_vp, C.c_int32, _f32p, _i32p],
    "nbx_batch_events": [_vp, _i64p],
    "nbx_batch_stream": [_vp],
    "nbx_lif_fwd_step": [C.POINTER(LifAdjParams), _vp, _vp, _vp, _vp, _vp],
    "nbx_lif_bwd_step": [C.POINTER(LifAdjParams), _vp, _vp, _vp, _vp, _vp, _vp],
    "nbx_lif_fwd_segment": [C.POINTER(LifAdjParams), _vp, _vp, _vp, _vp, _vp, _vp, C.c_int32, _vp],
    "nbx_lif_fwd_segment_sparse": [C.POINTER(LifAdjParams), _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int32, _vp],
    "nbx_lif_bwd_segment": [C.POINTER(LifAdjParams), _vp, _vp, _vp, _vp, _vp, _vp, C.c_int32, _vp],
    "nbx_readout_fwd": [C.c_int32, C.c_int64, C.c_float, _vp, _vp, _vp, _vp],
    "nbx_readout_bwd": [C.c_int32, C.c_int64, C.c_float, _vp, _vp, _vp, _vp],
}
_RESTYPES = {"nbx_last_error": C.c_char_p, "nbx_net_stream": _vp, "nbx_batch_stream": _vp}

_lib = None


def lib() -> C.CDLL:
    """Load libnbx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build the HIP library first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, C.c_int32)
        _lib = L
    return _lib


def call(name: str, *args) -> int:
    """Call a status-returning entry point; raise NbxError on failure."""
    L = lib()
    rc = getattr(L, name)(*args)
    if rc != NBX_OK:
        msg = L.nbx_last_error()
        raise NbxError(name, rc, msg.decode() if msg else "")
    return rc
