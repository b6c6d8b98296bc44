"""ctypes binding of libscarfhip.so (the C ABI declared in include/scarf_hip.h).

The library is the only compute path: if it is missing or fails to load, every
entry point raises ``RuntimeError``; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libscarfhip.so")

SH_OK = 0
SH_EINVAL = -1
SH_EHIP = -2
SH_EWORKSPACE = -3
SH_EUNSUPPORTED = -4

SH_F32, SH_BF16, SH_FP8, SH_F64 = 0, 1, 2, 3
DTYPES = {"fp32": SH_F32, "f32": SH_F32, "bf16": SH_BF16, "fp8": SH_FP8}

SH_FLAG_UNCERTIFIED = 1
SH_FLAG_SELF_MISSING = 2
SH_FLAG_SELF_NOT_FIRST = 4
SH_FLAG_THIN = 8

_p = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_int = ctypes.c_int
_f32 = ctypes.c_float
_f64 = ctypes.c_double
_sz = ctypes.c_size_t

# name -> (restype, argtypes)
_SIGS = {
    "sh_version": (_int, []),
    "sh_last_error": (ctypes.c_char_p, []),
    "sh_norm_stats": (_int, [_p, _p, _p, _i64, _i64, _p, _i32, _i32, _f64, _int, _p, _p, _p, _p, _p, _sz, _p]),
    "sh_norm_stats_workspace": (_sz, [_i64, _i64, _i32]),
    "sh_stats_merge": (_int, [_p, _p, _p, _i32, _i32, _p, _p, _p, _p]),
    "sh_densify": (_int, [_p, _p, _p, _p, _i64, _p, _i32, _i32, _f64, _int, _p, _p, _p, _f64, _int, _p, _i64, _p]),
    "sh_gram": (_int, [_p, _int, _i64, _i32, _i64, _int, _int, _p, _p, _p, _sz, _p]),
    "sh_gram_workspace": (_sz, [_i64, _i32, _int]),
    "sh_sparse_gram": (_int, [_p, _p, _p, _p, _i64, _i64, _p, _i32, _i32, _f64, _int, _p, _p,
                              _p, _f64, _int, _p, _p, _p, _sz, _p]),
    "sh_sparse_gram_workspace": (_sz, [_i64, _i64, _i32]),
    "sh_merge_matrix": (_int, [_p, _i32, _f64, _p, _p, _p, _i32, _p, _p, _p]),
    "sh_lanczos_workspace": (_sz, [_i32, _i32]),
    "sh_lanczos_steps": (_int, [_p, _i32, _p, _i32, _i32, _i32, _p, _p, _p, _p, _sz, _p]),
    "sh_dgemm": (_int, [_int, _i64, _i64, _i64, _f64, _p, _i64, _p, _i64, _f64, _p, _i64, _p]),
    "sh_symm_block_workspace": (_sz, [_i32, _i32, _i64]),
    "sh_symm_block": (_int, [_p, _i32, _p, _p, _i32, _i64, _f64, _f64, _f64, _p, _p, _sz, _p]),
    "sh_block_residual": (_int, [_p, _p, _i32, _i32, _i64, _p, _p, _p]),
    "sh_block_gram_workspace": (_sz, [_i32, _i32, _i32]),
    "sh_block_gram": (_int, [_p, _i64, _i32, _p, _i64, _i32, _i32, _p, _p, _sz, _p]),
    "sh_dense_z": (_int, [_p, _int, _i64, _i32, _i64, _p, _p, _f64, _p, _i64, _p]),
    "sh_dense_project": (_int, [_p, _int, _i64, _i32, _i64, _p, _p, _f64, _p, _i32, _p, _i64, _p]),
    "sh_project": (_int, [_p, _p, _p, _i64, _p, _i32, _i32, _f64, _int, _p, _p, _p, _f32, _p, _i64, _i32, _p, _p, _i64, _p]),
    "sh_knn_pack_k": (_i32, [_int, _i32]),
    "sh_knn_pack_row_bytes": (_i64, [_int, _i32]),
    "sh_knn_pack": (_int, [_p, _i64, _i32, _i64, _int, _int, _p, _p, _p]),
    "sh_knn_screen": (_int, [_p, _i64, _p, _i64, _i32, _int, _p, _p, _i32, _i32, _p, _p, _p, _p, _sz,
                            _p]),
    "sh_knn_screen_workspace": (_sz, [_i64, _i32, _i32]),
    "sh_knn_rerank": (_int, [_p, _i64, _i64, _p, _i64, _i64, _i32, _p, _p, _i32, _i32, _i32, _i64, _p, _f64, _f32, _f64, _p, _p, _p, _p]),
    "sh_knn_exact": (_int, [_p, _i64, _i64, _p, _i64, _p, _i64, _i64, _i32, _i32, _i64, _p, _p, _p, _p, _p, _p]),
    "sh_fix_knn_query": (_int, [_p, _p, _p, _i64, _i32, _p, _p, _p, _p]),
    "sh_locality_order_workspace": (_sz, [_i64]),
    "sh_locality_order": (_int, [_p, _i64, _i32, _i64, _p, _p, _sz, _p]),
    "sh_permute_rows": (_int, [_p, _i64, _p, _i64, _i64, _p, _i64, _int, _p]),
    "sh_smooth_knn": (_int, [_p, _p, _i64, _i32, _i64, _f32, _f32, _i32, _p, _p, _p, _p, _sz, _p]),
    "sh_smooth_knn_workspace": (_sz, [_i64, _i64]),
    "sh_mapping_score": (_int, [_p, _p, _i64, _i32, _p, _i32, _i64, _int, _f64, _f64, _int, _p, _p, _sz, _p]),
    "sh_mapping_score_workspace": (_sz, [_i64, _i32]),
    "sh_mapping_score_static": (_int, [_p, _p, _i64, _i32, _p, _i32, _i64, _int, _f64, _f64, _int, _p, _p]),
    "sh_mapping_score_static_max": (_sz, [_i32]),
    "sh_target_classes": (_int, [_p, _p, _i64, _i32, _p, _i64, _f64, _p, _p]),
    "sh_kmeans_pp_workspace": (_sz, [_i64, _i32]),
    "sh_kmeans_pp": (_int, [_p, _int, _i64, _i32, _i64, _p, _i32, _i32, _i64, _p, _p, _p, _sz, _p]),
    "sh_kmeans_assign_workspace": (_sz, [_i64, _i32]),
    "sh_kmeans_assign": (_int, [_p, _int, _i64, _i32, _i64, _p, _i32, _p, _p, _p, _sz, _p]),
    "sh_kmeans_update": (_int, [_p, _int, _i64, _i32, _i64, _p, _p, _i32, _p, _p, _p]),
    "sh_kmeans_gather": (_int, [_p, _int, _i32, _i64, _p, _p, _i64, _p, _p]),
    "sh_cov_from_gram": (_int, [_p, _i64, _p, _i64, _i32, _f64, _p, _p]),
    "sh_spd_sqrt_workspace": (_sz, [_i32]),
    "sh_spd_sqrt": (_int, [_p, _i32, _p, _p, _i32, _f64, _p, _p, _p, _sz, _p]),
    "sh_graph_csr_workspace": (_sz, [_i64, _i64, _i32, _i32]),
    "sh_graph_csr": (_int, [_p, _p, _i64, _i32, _i32, _i64, _int, _int, _p, _p, _p, _i64, _p, _p, _sz, _p]),
}

_lock = threading.Lock()
_lib = None


def exported_symbols():
    return [k for k in _SIGS]


def lib() -> ctypes.CDLL:
    """The loaded library; raises RuntimeError if it is not built or cannot load."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libscarfhip.so not found at {LIB_PATH}; build it with "
                "`python -m scarf_amd.build` (there is no CPU fallback)")
        try:
            L = ctypes.CDLL(LIB_PATH)
        except OSError as e:
            raise RuntimeError(f"failed to load {LIB_PATH}: {e}") from e
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
        return _lib


class ScarfHipError(RuntimeError):
    pass


def check(rc: int, what: str = ""):
    if rc == SH_OK:
        return
    msg = lib().sh_last_error().decode(errors="replace")
    if rc == SH_EINVAL:
        raise ValueError(f"{what}: {msg}")
    raise ScarfHipError(f"{what} failed ({rc}): {msg}")


def call(name: str, *args):
    rc = getattr(lib(), name)(*args)
    check(rc, name)
    return rc
