"""ctypes bindings of ``libabm_hip.so`` (the C ABI declared in include/abm_amd.h).

The library is built in-tree by ``abm_amd.build``.  Loading fails loudly when
it is missing: there is no CPU fallback anywhere in ``abm_amd``.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libabm_hip.so"
HEADER = PKG.parent / "include" / "abm_amd.h"

ABM_OK = 0
ABM_ERR_INVALID = -1
ABM_ERR_HIP = -2
ABM_ERR_UNSUPPORTED = -3

BOUNDARY = {"walls": 0, "infinite": 1}
ORDER_SEQUENTIAL = 0
ORDER_SYNCHRONOUS = 1
WANT_DIAG = 2
COUNT_RECHECKS = 4
VERIFY_FAST_PATH = 8
KAHAN_SUMS = 16

MODES = ["explore", "exploit", "relocate", "collide", "pool"]   # ABM_MODE_* order


class AbmError(RuntimeError):
    """A non-zero status from the C ABI (message from abm_last_error)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class AbmUnsupported(AbmError):
    pass


# ----------------------------------------------------------------- structs
class VFReplica(C.Structure):
    _fields_ = [
        ("n_agents", C.c_int32), ("n_rays", C.c_int32), ("boundary", C.c_int32),
        ("limit_movement", C.c_int32),
        ("fov0", C.c_double), ("fov1", C.c_double),
        ("width", C.c_double), ("height", C.c_double), ("window_pad", C.c_double),
        ("gam", C.c_double), ("v0", C.c_double), ("alp0", C.c_double), ("alp1", C.c_double),
        ("bet0", C.c_double), ("bet1", C.c_double), ("max_vel", C.c_double), ("max_th", C.c_double),
    ]


class BatchInfo(C.Structure):
    _fields_ = [
        ("n_replicas", C.c_int32), ("n_agents_total", C.c_int32), ("field_words", C.c_int64),
        ("max_agents", C.c_int32), ("max_rays", C.c_int32),
    ]


class VFStateF64(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("x", "y", "theta", "v", "radius", "alp0", "bet0", "v0")]


class VFStateF32(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("x", "y", "theta", "v", "radius", "alp0", "bet0", "v0")]


class ForageReplica(C.Structure):
    _fields_ = [
        ("n_agents", C.c_int32), ("n_rays", C.c_int32), ("visual_exclusion", C.c_int32),
        ("patchwise_exclusion", C.c_int32), ("tau", C.c_int32), ("pad_", C.c_int32),
        ("fov0", C.c_double), ("fov1", C.c_double),
        ("width", C.c_double), ("height", C.c_double), ("window_pad", C.c_double),
        ("vision_range", C.c_double),
        ("T_w", C.c_double), ("Eps_w", C.c_double), ("g_w", C.c_double), ("B_w", C.c_double),
        ("w_max", C.c_double),
        ("T_u", C.c_double), ("Eps_u", C.c_double), ("g_u", C.c_double), ("B_u", C.c_double),
        ("u_max", C.c_double),
        ("S_wu", C.c_double), ("S_uw", C.c_double), ("F_N", C.c_double), ("F_R", C.c_double),
        ("exp_vel_max", C.c_double), ("exp_theta_min", C.c_double), ("exp_theta_max", C.c_double),
        ("reloc_theta_max", C.c_double), ("exp_stop_ratio", C.c_double),
    ]


class ForageState(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "x", "y", "theta", "v", "w", "u", "I_priv", "collected_r", "collected_r_before",
        "novelty", "mode", "overriding_mode", "env_status", "exploited_patch_id", "radius")]


class ForageWorldParams(C.Structure):
    _fields_ = [
        ("n_patches", C.c_int32), ("patch_radius", C.c_int32), ("min_units", C.c_int32),
        ("max_units", C.c_int32), ("pooling_time", C.c_int32), ("regenerate", C.c_int32),
        ("teleport_exploit", C.c_int32), ("allow_border_overlap", C.c_int32),
        ("collide_agents", C.c_int32), ("ghost_mode", C.c_int32),
        ("consumption", C.c_double), ("min_quality", C.c_double), ("max_quality", C.c_double),
    ]


class PatchState(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "count", "id", "x", "y", "radius", "units", "left", "quality")]


WORLD_STEP, WORLD_RESUME, WORLD_FILL = 0, 1, 2                       # ABM_WORLD_* phases
WORLD_DONE, WORLD_NEED_WORDS, WORLD_TIMEOUT, WORLD_BAD_COUNT, WORLD_CROWDED = 0, 1, 2, 4, 5   # statuses
FORAGE_PATH_AUTO, FORAGE_PATH_SWEEP = 0, 1                             # ABM_FORAGE_PATH_*


class CSParams(C.Structure):
    _fields_ = [
        ("n_rays", C.c_int32), ("has_max_proj_size", C.c_int32),
        ("fov0", C.c_double), ("fov1", C.c_double), ("max_proj_size", C.c_double),
    ]


CS_ROWS, CS_SUM, CS_MAX = 0, 1, 2   # ABM_CS_* modes


class LogBuffers(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "posx", "posy", "orientation", "velocity", "w", "u", "I_priv", "collected_r", "mode",
        "expl_patch_id", "up", "down")]

_P = C.c_void_p
_I32 = C.c_int32
_D = C.c_double
_SIGS = {
    "abm_last_error": (C.c_char_p, []),
    "abm_build_info": (C.c_char_p, []),
    "abm_mt19937_peek": (C.c_int, [_P, _I32, _P, _P]),
    "abm_mt19937_skip": (C.c_int, [_P, _I32, _P]),
    "abm_mt19937_peek32": (C.c_int, [_P, _I32, _P, _P]),
    "abm_mt19937_skip32": (C.c_int, [_P, _I32, _P]),
    "abm_vf_batch_create": (C.c_int, [C.POINTER(VFReplica), _I32, C.POINTER(_P)]),
    "abm_vf_batch_destroy": (C.c_int, [_P]),
    "abm_vf_batch_get_info": (C.c_int, [_P, C.POINTER(BatchInfo)]),
    "abm_vf_batch_get_offsets": (C.c_int, [_P, _P, _P]),
    "abm_vf_step_f64": (C.c_int, [_P, C.POINTER(VFStateF64), _P, _P, C.c_uint32, _P]),
    "abm_vf_step_f32": (C.c_int, [_P, C.POINTER(VFStateF32), _P, _P, C.c_uint32, _P]),
    "abm_vf_f32_fallback_count": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                            C.POINTER(C.c_int64), _P]),
    "abm_vf_f64_near_ties": (C.c_int, [_P, _P, _P]),
    "abm_vf_set_launch": (C.c_int, [_P, _I32]),
    "abm_vf_last_kernels": (C.c_int, [_P, C.POINTER(C.c_uint32)]),
    "abm_vf_set_grouping": (C.c_int, [_P, _I32]),
    "abm_vf_autotune": (C.c_int, [_P, C.POINTER(VFStateF32), _I32, C.POINTER(_I32),
                                  C.POINTER(C.c_float), _P]),
    "abm_vf_projection_rows": (C.c_int, [_I32, _D, _D, _I32, _D, _D, _D, _D, _D, _D,
                                         _P, _P, _P, _I32, _P, _P, _P]),
    "abm_vf_state_variables": (C.c_int, [_I32, _I32, _P, _P, _P, _P, _P, _D, _D, _D, _P, _P, _P]),
    "abm_vf_state_variables_noalloc": (C.c_int, [_I32, _I32, _P, _P, _P, _P, _P, _D, _D, _D, _P, _P,
                                                 _P]),
    "abm_forage_batch_create": (C.c_int, [C.POINTER(ForageReplica), _I32, C.POINTER(_P)]),
    "abm_forage_batch_destroy": (C.c_int, [_P]),
    "abm_forage_batch_get_info": (C.c_int, [_P, C.POINTER(BatchInfo)]),
    "abm_forage_batch_get_offsets": (C.c_int, [_P, _P, _P]),
    "abm_forage_set_path": (C.c_int, [_P, _I32]),
    "abm_forage_sweep_counts": (C.c_int, [_P, _P, _P]),
    "abm_forage_step_f64": (C.c_int, [_P, C.POINTER(ForageState), _P, _P, _P, _P, _P, _P]),
    "abm_forage_projection": (C.c_int, [_I32, _D, _D, _I32, _D, _I32, _D, _D, _D, _D,
                                        _P, _P, _I32, _I32, _P, _P]),
    "abm_forage_world_create": (C.c_int, [_P, C.POINTER(ForageWorldParams), C.POINTER(_P)]),
    "abm_forage_world_destroy": (C.c_int, [_P]),
    "abm_forage_world_get_offsets": (C.c_int, [_P, _P]),
    "abm_forage_world_run": (C.c_int, [_P, _I32, C.POINTER(ForageState), _P, C.POINTER(PatchState),
                                       _P, _P, _P, _P, _P, _P, _P]),
    "abm_forage_world_run_dev": (C.c_int, [_P, _I32, C.POINTER(ForageState), _P, C.POINTER(PatchState),
                                           _P, _P, _P, _P, _P]),
    "abm_cs_projection": (C.c_int, [C.POINTER(CSParams), _I32, _P, _P, _P, _P, _P, _I32, _P, _P]),
    "abm_ensemble_stats": (C.c_int, [_I32, _P, _P, _P, _P, _I32, _I32, _D, _D, _P, _P]),
    "abm_log_record": (C.c_int, [C.POINTER(LogBuffers), C.c_int64, C.c_int64, _I32, _P, _P, _P, _P,
                                 _I32, _P, _P, _P, _P, _P, _P, _P, _I32, _I32, _P]),
}

_lib = None


def declared_symbols():
    """Function names declared in include/abm_amd.h."""
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(abm_\w+)\s*\(", text, re.M)))


def load():
    """Load the library once; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise AbmError(ABM_ERR_UNSUPPORTED,
                       f"{LIB_PATH} is not built; run `python -m abm_amd.build`")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int):
    """Raise AbmError for a non-zero status."""
    if rc != ABM_OK:
        msg = load().abm_last_error().decode(errors="replace")
        if rc == ABM_ERR_UNSUPPORTED:
            raise AbmUnsupported(rc, msg)
        raise AbmError(rc, msg)


def ptr(t):
    """Device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise AbmError(ABM_ERR_INVALID, "abm_amd needs device (ROCm) tensors")
    if not t.is_contiguous():
        raise AbmError(ABM_ERR_INVALID, "abm_amd needs contiguous tensors")
    return C.c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)
