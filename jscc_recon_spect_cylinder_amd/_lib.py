"""Loader of libjscc_hip.so and its C-ABI signatures (include/jscc_hip.h).

The library is the only compute path of this package: there is no CPU or
eager-PyTorch fallback.  If the library is missing, or no GPU is visible,
every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libjscc_hip.so")

_lib = None
_lock = threading.Lock()

c_int = ctypes.c_int
c_i64 = ctypes.c_int64
c_f32 = ctypes.c_float
c_f64 = ctypes.c_double
vp = ctypes.c_void_p

# name -> (restype, argtypes); mirrors include/jscc_hip.h
SIGNATURES = {
    "jscc_abi_version": (c_int, []),
    "jscc_last_error": (ctypes.c_char_p, []),
    "jscc_device_cu_count": (c_int, [vp]),
    "jscc_events_prepare": (c_int, [vp, c_i64, c_i64, vp, vp, vp, vp, c_i64, c_i64, c_i64, c_int, c_f32, c_f32,
                                    c_f32, c_f32, c_f32, c_f32, c_f32, c_f32, c_f64, c_f64, vp, vp, vp, vp]),
    "jscc_compact_workspace_bytes": (c_i64, [c_i64]),
    "jscc_compact_flags": (c_int, [vp, c_i64, vp, vp, vp, c_i64, vp]),
    "jscc_cone_rows": (c_int, [vp, vp, c_i64, vp, vp, vp, c_i64, vp, c_i64, vp, vp, vp, vp, c_int, c_f64, vp,
                               c_int, c_i64, vp, c_i64, vp, vp]),
    "jscc_cone_rows_counted": (c_int, [vp, vp, vp, c_i64, vp, vp, vp, c_i64, vp, c_i64, vp, vp, vp, vp, c_int, c_f64,
                                       vp, c_int, c_i64, vp, c_i64, vp, vp]),
    "jscc_cone_sens_grid": (c_int, [c_i64]),
    "jscc_cone_sens_workspace_bytes": (c_i64, [c_i64, c_int]),
    "jscc_cone_rows_sens": (c_int, [vp, vp, vp, c_i64, vp, vp, vp, c_i64, vp, c_i64, c_int, c_f64, vp, c_int, c_i64,
                                    vp, vp, c_f64, vp, vp, c_i64, vp]),
    "jscc_rows_keep_flags": (c_int, [vp, c_i64, c_f64, vp, vp, vp]),
    "jscc_rows_gather_normalize": (c_int, [vp, c_int, c_i64, c_i64, vp, c_i64, vp, c_int, vp, vp, c_i64, c_i64, vp]),
    "jscc_arows_normalize": (c_int, [vp, c_i64, c_i64, vp, c_i64, vp, c_i64, vp]),
    "jscc_quantize_fp8_rows": (c_int, [vp, c_int, c_i64, c_i64, c_i64, vp, vp, c_i64, vp, vp]),
    "jscc_quantize_rows": (c_int, [vp, c_int, c_i64, c_i64, c_i64, vp, vp, c_int, c_i64, vp, vp]),
    "jscc_rows_csr_map": (c_int, [vp, c_i64, c_i64, vp, vp, vp, c_i64, vp, c_i64, vp]),
    "jscc_materialize_rows": (c_int, [vp, c_i64, c_i64, c_i64, vp, vp, vp, c_i64, vp, c_i64, c_i64, vp, c_i64, c_int,
                                      vp, vp]),
    "jscc_proj_vec_width": (c_int, [c_int]),
    "jscc_proj_max_cols": (c_int, [c_int]),
    "jscc_proj_grid": (c_int, [c_i64]),
    "jscc_proj_workspace_bytes": (c_i64, [c_i64, c_i64]),
    "jscc_proj": (c_int, [c_int, vp, c_int, c_i64, c_i64, vp, c_i64, c_i64, vp, vp, vp, c_i64, c_i64, vp, c_i64, vp,
                          c_f32, c_int, vp, vp, c_i64, c_f32, c_int, c_int, vp, c_i64, vp, c_i64, vp, vp]),
    "jscc_proj_em": (c_int, [vp, c_int, c_i64, c_i64, vp, c_i64, c_i64, vp, vp, vp, c_i64, c_i64, vp, c_i64, c_f32,
                             c_int, vp, c_i64, c_f32, c_int, c_int, vp, c_i64, vp, vp, c_int, c_f32, c_f32, vp, c_f32,
                             c_int, vp, c_i64, vp, vp]),
    "jscc_proj_counted": (c_int, [c_int, vp, c_int, c_i64, c_i64, vp, vp, c_i64, c_int, c_int, c_i64, vp, vp, vp,
                                  c_i64, c_i64, vp, c_i64, c_f32, c_int, vp, c_i64, c_f32, c_int, c_int, vp, c_i64,
                                  vp, c_i64, vp, vp]),
    "jscc_proj_em_counted": (c_int, [vp, c_int, c_i64, c_i64, vp, vp, c_i64, c_int, c_int, c_i64, vp, vp, vp, c_i64,
                                     c_i64, vp, c_i64, c_f32, c_int, vp, c_i64, c_f32, c_int, c_int, vp, c_i64, vp, vp,
                                     c_int, c_f32, c_f32, vp, c_f32, c_int, vp, c_i64, vp, vp]),
    "jscc_proj_em_noalloc": (c_int, [vp, c_int, c_i64, c_i64, vp, c_i64, c_i64, vp, vp, vp, c_i64, c_i64, vp, c_i64,
                                     c_f32, c_int, vp, c_i64, c_f32, c_int, c_int, vp, c_i64, vp, vp, c_int, c_f32,
                                     c_f32, vp, c_f32, c_int, vp, vp]),
    "jscc_proj_joint_em": (c_int, [vp, c_int, c_i64, c_i64, vp, c_i64, c_i64, vp, vp, c_i64, c_int, c_f32,
                                   vp, c_i64, vp, c_i64, c_i64, vp, vp, c_i64, c_int, c_f32,
                                   vp, c_f32, c_int, c_int, vp, c_i64, vp, vp, c_int, c_f32, c_int, vp, c_i64, vp, vp]),
    "jscc_em_update": (c_int, [vp, vp, c_f32, vp, c_f32, vp, c_i64, c_int, c_f32, vp, vp]),
    "jscc_sensitivity_workspace_bytes": (c_i64, [c_i64, c_i64]),
    "jscc_sensitivity": (c_int, [vp, c_i64, c_i64, c_i64, vp, c_i64, c_i64, c_int, vp, vp, c_i64, vp, vp]),
    "jscc_fp8_selftest": (c_int, [vp, vp]),
    "jscc_cone_proj_grid": (c_int, [c_i64]),
    "jscc_cone_proj_workspace_bytes": (c_i64, [c_i64, c_i64]),
    "jscc_cone_proj": (c_int, [c_int, vp, vp, c_i64, vp, vp, vp, c_i64, vp, c_i64, c_int, c_f64, vp, vp, c_f32, c_int,
                               c_int, vp, vp, c_f32, c_int, vp, c_i64, vp, c_i64, vp, vp]),
    "jscc_sum_f64": (c_int, [vp, c_i64, c_i64, vp, vp]),
    "jscc_thin_counts": (c_int, [vp, c_i64, vp, c_i64, vp, c_i64, vp, vp, vp]),
    "jscc_osem_persistent_workspace_bytes": (c_i64, [c_i64]),
    "jscc_osem_persistent": (c_int, [vp, c_int, c_i64, c_i64, vp, vp, vp, c_int, c_int, vp, vp, c_f32, c_int, c_int,
                                     c_f32, vp, c_i64, vp, vp]),
    "jscc_cone_batch_groups": (c_int, [c_i64, c_int]),
    "jscc_cone_batch_workspace_bytes": (c_i64, [c_i64, c_int, c_i64, c_int]),
    "jscc_cone_proj_batched": (c_int, [c_int, vp, vp, c_i64, vp, c_int, vp, vp, vp, c_i64, vp, c_i64, c_int, c_f64,
                                       vp, c_i64, c_f32, c_int, c_int, vp, c_i64, c_int, c_f32, vp, c_i64, c_int,
                                       c_int, vp, c_i64, vp, vp]),
    "jscc_cone_batch_tiles": (c_int, [c_i64]),
    "jscc_cone_batch_tile_shape": (c_int, [c_i64, c_int, vp, vp, vp]),
    "jscc_coarse_weight_workspace_bytes": (c_i64, [c_i64, c_i64]),
    "jscc_coarse_weight": (c_int, [vp, c_i64, vp, c_i64, vp, vp, c_int, c_int, vp, vp, c_int, vp, vp, c_int, c_int,
                                    vp, c_i64, c_i64, vp, vp, c_i64, c_i64, c_f32, c_int, c_int, vp, c_i64, c_f32, c_int,
                                    vp, c_i64, vp, c_i64, vp, vp]),
    "jscc_frames_transpose_bf16": (c_int, [vp, c_i64, c_i64, c_i64, vp, c_i64, vp]),
    "jscc_frames_split": (c_int, [vp, c_i64, c_i64, c_int, c_i64, vp, vp, c_i64, vp]),
    "jscc_frames_split_count": (c_int, [c_i64, c_i64, c_int]),
    "jscc_frames_partial_bytes": (c_i64, [c_i64, c_int, c_int]),
    "jscc_frames_gemm": (c_int, [vp, c_i64, c_i64, c_i64, vp, vp, c_i64, c_int, c_int, vp, c_i64, vp]),
    "jscc_frames_ratio": (c_int, [vp, c_int, c_i64, c_i64, c_i64, c_int, vp, c_i64, c_f32, c_int, vp, vp, c_i64, vp,
                                  vp]),
    "jscc_frames_update": (c_int, [vp, c_int, c_i64, c_int, vp, vp, c_i64, c_int, c_f32, vp, vp, c_i64, vp]),
    "jscc_prof_enable": (c_int, [c_int]),
    "jscc_prof_read": (c_int, [c_int, vp, vp]),
}

PROF_CONE, PROF_PROJ, PROF_PROJ_SUM, PROF_CONE_PROJ, PROF_COARSE, PROF_FRAMES = 0, 1, 2, 3, 4, 5


class JsccError(RuntimeError):
    """A libjscc_hip entry point returned a non-zero status."""


def load(path: str | None = None):
    """Load (once) and return the ctypes handle of libjscc_hip.so."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise JsccError(
                f"{p} is missing: build the gfx950 kernels first "
                "(python -c 'import __graft_entry__ as g; g.build()' or "
                "python jscc_recon_spect_cylinder_amd/build.py)")
        lib = ctypes.CDLL(p)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def lib():
    return _lib if _lib is not None else load()


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().jscc_last_error()
        raise JsccError(f"{what}: {msg.decode() if msg else 'error'} (status {rc})")


def exported_symbols() -> list[str]:
    return list(SIGNATURES)
