"""ctypes binding of libeplace.so (include/eplace.h).

The library is the only compute path: loading fails loudly when it is
missing, and every placement call needs a HIP device.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libeplace.so"

EP_OK = 0
EP_PENDING = 1
EP_ERR_INVALID_ARG = -1
EP_ERR_NO_BELS = -2
EP_ERR_PLACEMENT_FAILED = -3
EP_ERR_OOM = -4
EP_ERR_ADMISSION_REFUSED = -5
EP_ERR_DEVICE = -6
EP_ERR_NONFINITE = -7
EP_ERR_NOT_FOUND = -8
EP_ERR_CANCELLED = -9

EP_PREC_FP64 = 0
EP_PREC_FAST = 1
EP_FIELD_REFERENCE = 0
EP_FIELD_IDXST = 1

_p = C.c_void_p


class EpDesign(C.Structure):
    _fields_ = [
        ("n_mov", C.c_int32), ("n_fixed", C.c_int32), ("n_nets", C.c_int32),
        ("W", C.c_int32), ("H", C.c_int32), ("Bx", C.c_int32), ("By", C.c_int32),
        ("reserved", C.c_int32), ("n_pins", C.c_int64),
        ("x0", _p), ("y0", _p), ("area", _p), ("fixed_x", _p), ("fixed_y", _p),
        ("net_off", _p), ("pin_cell", _p), ("pinw", _p), ("net_weight", _p),
    ]


class EpConfig(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("wl_coeff", C.c_double), ("target_util", C.c_double),
        ("timing_driven", C.c_int32), ("max_iters", C.c_int32),
        ("target_density", C.c_double), ("timing_weight", C.c_double),
        ("nesterov_step_size", C.c_double), ("legalize_interval", C.c_int32),
        ("min_iters", C.c_int32), ("overflow_target", C.c_double),
        ("density_eta", C.c_double), ("grow_mult", C.c_double),
        ("grow_threshold", C.c_double), ("grow_add", C.c_double),
        ("stop_every_iter", C.c_int32), ("log_hpwl", C.c_int32),
        ("check_nonfinite", C.c_int32), ("field_mode", C.c_int32),
    ]


class EpIterRec(C.Structure):
    _fields_ = [
        ("iter", C.c_int32), ("checked", C.c_int32), ("overflow", C.c_double),
        ("lambda_", C.c_double), ("step_size_x", C.c_double), ("step_size_y", C.c_double),
        ("step_norm_x", C.c_double), ("step_norm_y", C.c_double), ("hpwl", C.c_double),
        ("wl_norm", C.c_double), ("den_norm", C.c_double),
    ]


class EpResult(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("iters", C.c_int32), ("converged", C.c_int32),
        ("precision", C.c_int32), ("overflow", C.c_double), ("hpwl", C.c_double),
        ("lambda_", C.c_double), ("t_submit_ms", C.c_double), ("t_admit_ms", C.c_double),
        ("t_done_ms", C.c_double),
    ]


class EpLegalProblem(C.Structure):
    _fields_ = [("n_bels", C.c_int32), ("n_children", C.c_int32), ("bel_x", _p), ("bel_y", _p),
                ("bel_bucket", _p), ("bel_taken", _p), ("cell_bucket", _p), ("child_root", _p),
                ("child_dx", _p), ("child_dy", _p), ("child_bucket", _p), ("child_fixed", _p),
                ("net_has_driver", _p)]


class EpLegalResult(C.Structure):
    _fields_ = [("best_iter", C.c_int32), ("legalizations", C.c_int32), ("best_hpwl", C.c_double),
                ("final_displacement", C.c_double)]


class EpServeConfig(C.Structure):
    _fields_ = [("max_slots", C.c_uint32), ("iters_per_sync", C.c_uint32),
                ("hbm_budget_bytes", C.c_uint64)]


class EpServeStats(C.Structure):
    _fields_ = [("admitted", C.c_uint64), ("finished", C.c_uint64), ("batch_builds", C.c_uint64),
                ("graph_captures", C.c_uint64), ("graph_updates", C.c_uint64), ("chunks", C.c_uint64),
                ("failed", C.c_int32),
                ("reserved", C.c_int32)]


class EpPhaseTime(C.Structure):
    _fields_ = [("name", C.c_char * 24), ("launches", C.c_int32), ("reserved", C.c_int32),
                ("avg_ms", C.c_double)]


class EpGenParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("n_cells", C.c_int32), ("n_fixed", C.c_int32),
        ("n_nets", C.c_int32), ("W", C.c_int32), ("H", C.c_int32), ("Bx", C.c_int32),
        ("By", C.c_int32), ("n_clusters", C.c_int32), ("area_mix", C.c_int32),
        ("n_types", C.c_int32), ("n_high_fanout", C.c_int32), ("hf_min", C.c_int32),
        ("hf_max", C.c_int32), ("locality", C.c_double), ("type1_frac", C.c_double),
    ]


class EpGenDesign(C.Structure):
    _fields_ = [
        ("d", EpDesign), ("n_types", C.c_int32), ("n_bels", C.c_int32),
        ("cell_type", _p), ("bel_x", _p), ("bel_y", _p), ("bel_type", _p), ("init_bel", _p),
    ]


# Every entry point the header declares, with its ctypes signature.
SIGNATURES = {
    "ep_version": (C.c_int, []),
    "ep_create": (C.c_int, [C.POINTER(_p), C.c_int, C.c_uint64]),
    "ep_destroy": (None, [_p]),
    "ep_last_error": (C.c_char_p, [_p]),
    "ep_config_default": (None, [C.POINTER(EpConfig)]),
    "ep_design_upload": (C.c_int, [_p, C.POINTER(EpDesign), C.c_int32, C.POINTER(C.c_uint32)]),
    "ep_design_free": (C.c_int, [_p, C.c_uint32]),
    "ep_design_bytes": (C.c_uint64, [C.POINTER(EpDesign), C.c_int32]),
    "ep_place": (C.c_int, [_p, C.c_uint32, C.POINTER(EpConfig), C.POINTER(EpResult)]),
    "ep_place_noalloc": (C.c_int, [_p, C.c_uint32, C.POINTER(EpConfig), C.POINTER(EpResult)]),
    "ep_place_batch": (C.c_int, [_p, C.POINTER(C.c_uint32), C.c_uint32, C.POINTER(EpConfig),
                                 C.POINTER(EpResult)]),
    "ep_positions": (C.c_int, [_p, C.c_uint32, _p, _p]),
    "ep_iter_log": (C.c_int, [_p, C.c_uint32, C.POINTER(EpIterRec), C.c_uint32,
                              C.POINTER(C.c_uint32)]),
    "ep_eval_config": (C.c_int, [_p, C.POINTER(EpConfig)]),
    "ep_eval_wa_grad": (C.c_int, [_p, C.c_uint32, _p, _p, _p, _p]),
    "ep_eval_density": (C.c_int, [_p, C.c_uint32, _p, _p, _p, _p]),
    "ep_eval_bin_ranges": (C.c_int, [_p, C.c_uint32, _p, _p, _p]),
    "ep_eval_field": (C.c_int, [_p, C.c_uint32, _p, _p, _p]),
    "ep_eval_density_grad": (C.c_int, [_p, C.c_uint32, _p, _p, _p, _p, _p, _p]),
    "ep_eval_hpwl": (C.c_int, [_p, C.c_uint32, _p, _p, _p]),
    "ep_serve_start": (C.c_int, [_p, C.POINTER(EpServeConfig)]),
    "ep_serve_submit": (C.c_int, [_p, C.POINTER(EpDesign), C.POINTER(EpConfig),
                                  C.POINTER(C.c_uint64)]),
    "ep_serve_poll": (C.c_int, [_p, C.c_uint64, C.POINTER(EpResult)]),
    "ep_serve_positions": (C.c_int, [_p, C.c_uint64, _p, _p]),
    "ep_serve_release": (C.c_int, [_p, C.c_uint64]),
    "ep_serve_cancel": (C.c_int, [_p, C.c_uint64]),
    "ep_serve_stats_get": (C.c_int, [_p, C.POINTER(EpServeStats)]),
    "ep_serve_stop": (C.c_int, [_p]),
    "ep_profile_iteration": (C.c_int, [_p, C.POINTER(C.c_uint32), C.c_uint32, C.POINTER(EpConfig),
                                       C.c_uint32, C.c_uint32, C.POINTER(EpPhaseTime), C.c_uint32,
                                       C.POINTER(C.c_uint32)]),
    "ep_calibrate_copy": (C.c_int, [_p, C.c_uint64]),
    "ep_autotune": (C.c_int, [_p, C.POINTER(C.c_uint32), C.c_uint32, C.c_uint32, C.c_char_p]),
    "ep_tune_table": (C.c_int, [_p, C.POINTER(C.c_int32), C.c_uint32, C.POINTER(C.c_uint32)]),
    "ep_tune_set": (C.c_int, [_p, C.c_int32, C.c_int32, C.c_int32]),
    "ep_legalize": (C.c_int, [C.c_int32, _p, _p, _p, C.c_int32, _p, _p, _p, _p, _p, _p]),
    "ep_legalize_problem": (C.c_int, [C.POINTER(EpLegalProblem), C.c_int32, _p, _p, _p, _p, _p]),
    "ep_legal_hpwl": (C.c_int, [C.POINTER(EpDesign), C.POINTER(EpLegalProblem), _p, _p, _p]),
    "ep_place_legalized": (C.c_int, [_p, C.c_uint32, C.POINTER(EpConfig), C.POINTER(EpLegalProblem),
                                     C.POINTER(EpResult), C.POINTER(EpLegalResult), _p, _p]),
    "ep_generate": (C.c_int, [C.POINTER(EpGenParams), C.POINTER(C.POINTER(EpGenDesign))]),
    "ep_gen_free": (None, [C.POINTER(EpGenDesign)]),
}

_LIB = None


def lib() -> C.CDLL:
    """The loaded libeplace.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        path = Path(os.environ.get("EP_LIB", LIB_PATH))
        if not path.exists():
            raise RuntimeError(
                f"{path} is missing: build it with `python -m eisenjoch_amd.build` "
                "(there is no CPU fallback)")
        L = C.CDLL(str(path))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(_p)
