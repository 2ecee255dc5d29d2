"""ctypes binding of libkdigest_hip.so (the C ABI in include/kdigest_hip.h).

The library is the only compute path: importing this module never falls back
to a CPU implementation.  If the shared object is missing the import raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# KDIGEST_AMD_LIB names another build of the library (an experiment or an
# installed copy); by default the in-tree build.
LIB_PATH = os.environ.get("KDIGEST_AMD_LIB") or os.path.join(_HERE, "libkdigest_hip.so")

KD_OK, KD_EINVAL, KD_EEMPTY, KD_ERANGE, KD_ESERDE, KD_ETRUNC, KD_EDEVICE, KD_ENOMEM = (
    0, -1, -2, -3, -4, -5, -6, -7)

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_u32p = C.POINTER(C.c_uint32)
_u8p = C.POINTER(C.c_uint8)
_vp = C.c_void_p
_i64 = C.c_int64
_d = C.c_double
_i = C.c_int

# name -> (restype, argtypes); the exported set of include/kdigest_hip.h.
SIGNATURES = {
    "kd_last_error": (C.c_char_p, []),
    "kd_session_create": (_i, [_i, C.POINTER(_vp)]),
    "kd_session_destroy": (None, [_vp]),
    "kd_session_sync": (_i, [_vp]),
    "kd_session_profile": (_i, [_vp, _i]),
    "kd_session_set_select": (_i, [_vp, _i]),
    "kd_session_set_fused": (_i, [_vp, _i]),
    "kd_session_set_partition_shape": (_i, [_vp, _i, _i64]),
    "kd_session_set_select_shape": (_i, [_vp, _i64, _i64]),
    "kd_session_set_overlap": (_i, [_vp, _i]),
    "kd_session_set_heavy_keys": (_i, [_vp, _i64]),
    "kd_session_set_radix_key_bits": (_i, [_vp, _i]),
    "kd_session_autotune_grouped": (_i, [_vp, _i, _i64, _d, _vp, _vp, _i64, _i, _i64p, _dp]),
    "kd_autotune_save": (_i, [C.c_char_p]),
    "kd_session_shape": (_i, [_vp, _i64p]),
    "kd_autotune_load": (_i, [C.c_char_p, _i64p]),
    "kd_session_autotune_grouped_host": (_i, [_vp, _i, _i64, _d, _u32p, _dp, _i64, _i, _i64p, _dp]),
    "kd_session_path_counts": (_i, [_vp, _i64p]),
    "kd_session_phase_stats": (_i64, [_vp, C.c_char_p, _i64]),
    "kd_compress": (_i, [_vp, _i, _d, _d, _dp, _dp, _d, _i64, _i64, _dp, _dp, _i64p]),
    "kd_quantile": (_i, [_vp, _i64, _i64p, _dp, _dp, _dp, _i64p, _dp, _i64, _dp]),
    "kd_cdf": (_i, [_vp, _i64, _i64p, _dp, _dp, _dp, _i64p, _dp, _i64, _dp]),
    "kd_scale_eval": (_i, [_vp, _i, _d, _dp, _i64, _dp]),
    "kd_math_eval": (_i, [_vp, _i, _dp, _i64, _dp]),
    "kd_scale_normalizer": (_d, [_i, _d, _d]),
    "kd_scale_k": (_d, [_i, _d, _d]),
    "kd_add_repeated": (_d, [_d, _d, _i64]),
    "kd_add_sequence": (_i64, [_dp, _dp, _i64, _i, _d, _dp, _dp]),
    "kd_scale_q": (_d, [_i, _d, _d]),
    "kd_tdigest_create": (_i, [_vp, _i, _d, _d, C.POINTER(_vp)]),
    "kd_tdigest_destroy": (None, [_vp]),
    "kd_tdigest_add": (_i, [_vp, _d, _d]),
    "kd_tdigest_add_at": (_i, [_vp, _d, _d, _d]),
    "kd_tdigest_add_batch": (_i, [_vp, _dp, _i64, _i]),
    "kd_tdigest_add_batch_device": (_i, [_vp, _vp, _i64]),
    "kd_tdigest_add_batches_device": (_i, [_vp, _vp, _i64, _i64]),
    "kd_tdigest_merge": (_i, [_vp]),
    "kd_tdigest_quantile": (_i, [_vp, _dp, _i64, _dp]),
    "kd_tdigest_cdf": (_i, [_vp, _dp, _i64, _dp]),
    "kd_tdigest_merge_into": (_i, [_vp, _vp]),
    "kd_tdigest_to_bytes": (_i64, [_vp, _u8p, _i64]),
    "kd_tdigest_from_bytes": (_i, [_vp, _u8p, _i64, _i, C.POINTER(_vp)]),
    "kd_tdigest_from_centroids": (_i, [_vp, _i, _d, _dp, _dp, _i64, _d, _d, C.POINTER(_vp)]),
    "kd_tdigest_size": (_i64, [_vp]),
    "kd_tdigest_centroids": (_i, [_vp, _dp, _dp, _i64, _i64p]),
    "kd_tdigest_state": (_i, [_vp, _dp]),
    "kd_grouped_create": (_i, [_vp, _i, _i64, _d, _d, C.POINTER(_vp)]),
    "kd_grouped_destroy": (None, [_vp]),
    "kd_grouped_add_batch": (_i, [_vp, _u32p, _dp, _i64]),
    "kd_grouped_add_batch_device": (_i, [_vp, _vp, _vp, _i64]),
    "kd_grouped_quantile": (_i, [_vp, _i64p, _dp, _i64, _dp]),
    "kd_grouped_cdf": (_i, [_vp, _i64p, _dp, _i64, _dp]),
    "kd_grouped_quantile_device": (_i, [_vp, _vp, _vp, _i64, _vp]),
    "kd_grouped_quantile_device_noalloc": (_i, [_vp, _vp, _vp, _i64, _vp]),
    "kd_grouped_to_bytes": (_i64, [_vp, _i64, _u8p, _i64]),
    "kd_grouped_wire_size": (_i, [_vp, _i64p]),
    "kd_grouped_params": (_i, [_vp, _dp, _dp]),
    "kd_batch_params": (_i, [_vp, _dp, _dp]),
    "kd_keymap_create": (_i, [_vp, _i64, C.POINTER(_vp)]),
    "kd_keymap_destroy": (None, [_vp]),
    "kd_keymap_map_device": (_i, [_vp, _vp, _i64, _vp, _i]),
    "kd_keymap_map": (_i, [_vp, C.POINTER(C.c_uint32), _i64, C.POINTER(C.c_uint32), _i]),
    "kd_keymap_size": (_i, [_vp, _i64p]),
    "kd_keymap_keys": (_i, [_vp, C.POINTER(C.c_uint32), _i64]),
    "kd_batch_from_bytes_device": (_i, [_vp, _i, _i64, _vp, _i64, _vp, C.POINTER(_vp)]),
    "kd_batch_from_bytes_all": (_i, [_vp, _i, _i64, _u8p, _i64, _i64p, C.POINTER(_vp)]),
    "kd_grouped_to_bytes_device": (_i, [_vp, _vp, _i64, _vp]),
    "kd_grouped_to_bytes_all": (_i, [_vp, _u8p, _i64, _i64p]),
    "kd_grouped_from_bytes_device": (_i, [_vp, _i, _i64, _vp, _i64, _vp, C.POINTER(_vp)]),
    "kd_grouped_from_bytes_all": (_i, [_vp, _i, _i64, _u8p, _i64, _i64p, C.POINTER(_vp)]),
    "kd_grouped_centroid_count": (_i, [_vp, _i64, _i64p]),
    "kd_grouped_total_weight": (_i, [_vp, _i64, _dp]),
    "kd_grouped_centroids": (_i, [_vp, _i64, _dp, _dp, _i64, _i64p]),
    "kd_grouped_key_state": (_i, [_vp, _i64, _dp]),
    "kd_batch_create": (_i, [_vp, _i, _i64, _d, _d, C.POINTER(_vp)]),
    "kd_batch_destroy": (None, [_vp]),
    "kd_batch_add_at_device": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "kd_batch_add_at": (_i, [_vp, _i64p, _dp, _dp, _dp]),
    "kd_batch_flush": (_i, [_vp]),
    "kd_batch_merge_tree": (_i, [_vp]),
    "kd_batch_set_graph": (_i, [_vp, _i]),
    "kd_batch_state": (_i, [_vp, _i64, _dp]),
    "kd_batch_centroids": (_i, [_vp, _i64, _dp, _dp, _i64, _i64p]),
    "kd_batch_to_tdigest": (_i, [_vp, _i64, C.POINTER(_vp)]),
    "kd_server_create": (_i, [_vp, _i, _i64, _d, _d, _i64, _i64, C.POINTER(_vp)]),
    "kd_server_destroy": (None, [_vp]),
    "kd_server_submit": (_i, [_vp, _i, _i64, _dp, _i64, _i64p]),
    "kd_server_submit_many": (_i, [_vp, _i64, C.POINTER(_i), _i64p, _i64p, _dp, _i64p]),
    "kd_server_take_completed": (_i, [_vp, _i64p, _dp, _i64, _i64p]),
    "kd_server_clock": (_i, [_vp, _dp]),
    "kd_server_step": (_i, [_vp, _i, _i64p]),
    "kd_server_start": (_i, [_vp, _i]),
    "kd_server_stop": (_i, [_vp]),
    "kd_server_result": (_i, [_vp, _i64, _dp, _i64, C.POINTER(_i)]),
    "kd_server_pending": (_i, [_vp, _i64p]),
    "kd_server_stats": (_i, [_vp, _i64p]),
    "kd_server_key_state": (_i, [_vp, _i64, _dp]),
    "kd_server_key_centroids": (_i, [_vp, _i64, _dp, _dp, _i64, _i64p]),
}

_lib: C.CDLL | None = None


def load() -> C.CDLL:
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m kdigest_amd.build_ext` "
            "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


class KdigestError(Exception):
    """A failed call into the HIP library (message from kd_last_error)."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


def check(rc: int) -> None:
    if rc != KD_OK:
        msg = load().kd_last_error().decode(errors="replace")
        raise KdigestError(rc, msg)
