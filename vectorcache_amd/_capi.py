"""ctypes bindings of libvectorcache_hip.so (include/vectorcache_hip.h).

Importing this module loads the in-tree HIP library and raises ImportError if
it is not built: there is no CPU fallback on the product path.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libvectorcache_hip.so"

if not LIB_PATH.exists():
    raise ImportError(f"{LIB_PATH} is not built; run `python -m vectorcache_amd.build` (HIP, gfx950)")

# torch bundles its own HIP runtime under the same soname (libamdhip64.so.7).
# Loaded first, it is the one runtime of the process and the library binds to
# it; loaded after this library, torch would bring a second runtime that then
# finds no device. So torch, when present, is imported before the library.
try:
    import torch  # noqa: F401
except ImportError:
    pass

lib = ctypes.CDLL(str(LIB_PATH))

P = ctypes.c_void_p
S = ctypes.c_size_t
F = ctypes.c_float
I = ctypes.c_int

VC_OK, VC_ERR_INVALID_ARGUMENT, VC_ERR_STATE, VC_ERR_OUT_OF_RANGE, VC_ERR_DEVICE, VC_ERR_OOM = range(6)


class VectorCacheError(RuntimeError):
    """Device or allocation failure (VC_ERR_DEVICE, VC_ERR_OOM)."""


class DeviceOutOfMemory(VectorCacheError):
    pass


class IndexStateError(RuntimeError):
    """The reference's std::runtime_error for call-order errors (search before
    prepare) and for file / format errors (its vectorcache::Error)."""


class BucketParams(ctypes.Structure):
    """vc_bucket_params."""

    _fields_ = [("n_probe", ctypes.c_size_t), ("max_bucket_size", ctypes.c_size_t), ("energy_soft", ctypes.c_size_t),
                ("expected_n", ctypes.c_size_t), ("kmeans_iters", ctypes.c_size_t), ("kmeans_sample", ctypes.c_size_t),
                ("seed", ctypes.c_uint64), ("residual", ctypes.c_int)]


class ServerParams(ctypes.Structure):
    """vc_server_params."""

    _fields_ = [("slots", ctypes.c_size_t), ("max_batch_queries", ctypes.c_size_t), ("max_k", ctypes.c_size_t),
                ("hbm_budget_bytes", ctypes.c_size_t), ("reserve_vectors", ctypes.c_size_t),
                ("max_vectors", ctypes.c_size_t), ("add_chunk", ctypes.c_size_t)]


class RequestInfo(ctypes.Structure):
    """vc_request_info."""

    _fields_ = [("t_submit", ctypes.c_double), ("t_start", ctypes.c_double), ("t_done", ctypes.c_double),
                ("n_snapshot", ctypes.c_size_t), ("k_eff", ctypes.c_size_t), ("batch_queries", ctypes.c_size_t),
                ("batch", ctypes.c_uint64), ("slot", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    """vc_server_stats."""

    _fields_ = [("requests", ctypes.c_uint64), ("queries", ctypes.c_uint64), ("batches", ctypes.c_uint64),
                ("adds", ctypes.c_uint64), ("vectors_added", ctypes.c_uint64), ("admission_waits", ctypes.c_uint64),
                ("in_flight_peak", ctypes.c_size_t), ("pool_peak_bytes", ctypes.c_size_t),
                ("pool_bytes", ctypes.c_size_t), ("size", ctypes.c_size_t)]


class MultiPlanInfo(ctypes.Structure):
    """vc_multi_plan_info."""

    _fields_ = [("tile", ctypes.c_size_t), ("vec_per_thread", ctypes.c_size_t), ("item_scale", ctypes.c_float),
                ("workgroups_per_cu", ctypes.c_size_t), ("items", ctypes.c_size_t), ("grid", ctypes.c_size_t),
                ("lds_bytes", ctypes.c_size_t), ("measured_ms", ctypes.c_double)]


class ScanChoice(ctypes.Structure):
    """vc_scan_choice."""

    _fields_ = [("tile", ctypes.c_size_t), ("vec_per_thread", ctypes.c_size_t), ("range_scale", ctypes.c_float),
                ("ms", ctypes.c_double), ("rule_ms", ctypes.c_double), ("rule_tile", ctypes.c_size_t)]


class VcexKey(ctypes.Structure):
    """vc_vcex_key: the key of a VCEX exact top-k cache file."""

    _fields_ = [("source_label", ctypes.c_char_p), ("split", ctypes.c_char_p), ("query_split", ctypes.c_char_p),
                ("index_n", ctypes.c_uint64), ("query_n", ctypes.c_uint64), ("k", ctypes.c_uint64),
                ("seed", ctypes.c_uint64)]


_SIGS = {
    "vc_last_error": ([], ctypes.c_char_p),
    "vc_version": ([], ctypes.c_char_p),
    "vc_index_create": ([S, S, ctypes.POINTER(P)], I),
    "vc_index_destroy": ([P], I),
    "vc_index_calibrate": ([P, P, S], I),
    "vc_index_calibrate_device": ([P, P, S], I),
    "vc_index_add": ([P, P, S], I),
    "vc_index_add_device": ([P, P, S, P], I),
    "vc_index_prepare": ([P], I),
    "vc_index_search": ([P, P, S, S, P, P, ctypes.POINTER(S)], I),
    "vc_index_search_device": ([P, P, S, S, P, P, ctypes.POINTER(S), P], I),
    "vc_index_search_device_noalloc": ([P, P, S, S, P, P, ctypes.POINTER(S), P], I),
    "vc_index_search_prefix_device": ([P, P, S, S, S, P, P, ctypes.POINTER(S), P], I),
    "vc_multi_create": ([P, S, P, P, ctypes.POINTER(P)], I),
    "vc_multi_destroy": ([P], I),
    "vc_multi_rows": ([P, S, ctypes.POINTER(P), ctypes.POINTER(P), ctypes.POINTER(P), ctypes.POINTER(S)], I),
    "vc_multi_set_queries": ([P, S, P, P], I),
    "vc_multi_get_results": ([P, S, P, P, P], I),
    "vc_multi_search": ([P, P], I),
    "vc_multi_capture": ([P, P], I),
    "vc_multi_replay": ([P, P], I),
    "vc_multi_set_plan": ([P, S, S, F, S], I),
    "vc_multi_get_plan": ([P, ctypes.POINTER(MultiPlanInfo)], I),
    "vc_multi_collection_plan": ([P, S, ctypes.POINTER(S), ctypes.POINTER(S)], I),
    "vc_multi_tune": ([P, S, ctypes.POINTER(MultiPlanInfo)], I),
    "vc_tune_scan": ([S, S, S, S, S, S, ctypes.POINTER(ScanChoice)], I),
    "vc_tune_save": ([ctypes.c_char_p], I),
    "vc_tune_load": ([ctypes.c_char_p, ctypes.POINTER(S)], I),
    "vc_tune_clear": ([], I),
    "vc_tune_count": ([], S),
    "vc_server_params_default": ([ctypes.POINTER(ServerParams)], I),
    "vc_server_batch_bytes": ([P, S, S, S], S),
    "vc_server_create": ([P, ctypes.POINTER(ServerParams), ctypes.POINTER(P)], I),
    "vc_server_destroy": ([P], I),
    "vc_server_submit_search": ([P, P, S, S, ctypes.POINTER(ctypes.c_uint64)], I),
    "vc_server_submit_add": ([P, P, S, ctypes.POINTER(ctypes.c_uint64)], I),
    "vc_server_wait": ([P, ctypes.c_uint64, P, P, ctypes.POINTER(RequestInfo)], I),
    "vc_server_flush": ([P], I),
    "vc_server_get_stats": ([P, ctypes.POINTER(ServerStats)], I),
    "vc_noalloc_scratch_bytes": ([], S),
    "vc_index_dim": ([P], S),
    "vc_index_size": ([P], S),
    "vc_index_bits": ([P], S),
    "vc_index_is_prepared": ([P], I),
    "vc_index_is_calibrated": ([P], I),
    "vc_index_codes_bytes": ([P], S),
    "vc_index_index_bytes": ([P], S),
    "vc_index_export_packed": ([P, P], I),
    "vc_index_export_flat": ([P, P], I),
    "vc_index_export_perm0": ([P, P], I),
    "vc_index_export_seq": ([P, P], I),
    "vc_index_export_alpha": ([P, P], I),
    "vc_index_export_gpu32": ([P, P], I),
    "vc_gpu32_words": ([S, S, S], S),
    "vc_score_perm0": ([P, S, S, S, S, P, P, S, P, P, P, P, S, P, P, P, P, ctypes.POINTER(S)], I),
    "vc_gpu32_to_perm0": ([P, S, S, S, P], I),
    "vc_index_get_calibration": ([P, P, P], I),
    "vc_index_prepare_queries": ([P, P, S, P, P, P, P], I),
    "vc_rotation_tables": ([S, P, P, P, ctypes.POINTER(S)], I),
    "vc_codebook": ([S, S, P, P], I),
    "vc_encode": ([P, S, S, S, P, P, P, P], I),
    "vc_fit_calibration": ([P, S, S, S, P, P], I),
    "vc_rotate": ([P, S, S, P], I),
    "vc_search_perm0": ([P, S, S, S, S, P, P, S, P, P, P, P, P, ctypes.POINTER(S)], I),
    "vc_merge_results": ([S, S, P, P, S, P, P], I),
    "vc_score_perm0_into": ([P, S, S, S, S, P, P, S, P, P, P, P, S, P, P, P, P, P, P], I),
    "vc_heaps_to_results": ([S, S, P, P, P, P, P], I),
    "vc_prepare_queries": ([P, S, S, S, P, P, ctypes.POINTER(P)], I),
    "vc_prepare_queries_device": ([P, S, S, S, P, P, P, ctypes.POINTER(P)], I),
    "vc_prepared_destroy": ([P], I),
    "vc_prepared_shape": ([P, ctypes.POINTER(S), ctypes.POINTER(S), ctypes.POINTER(S), ctypes.POINTER(S)], I),
    "vc_prepared_export": ([P, P, P, P, P], I),
    "vc_prepared_query_at": ([P, S, ctypes.POINTER(P)], I),
    "vc_score_prepared": ([P, S, P, P, S, P, P, S, P, P, P, P, ctypes.POINTER(S)], I),
    "vc_score_prepared_into": ([P, S, P, P, S, P, P, S, P, P, P, P, P, P], I),
    "vc_codes_flat_from_packed": ([P, S, S, S, P], I),
    "vc_repack_perm0": ([P, S, S, S, P], I),
    "vc_repack_seq": ([P, S, S, S, P], I),
    "vc_float_rescore": ([S, S, P, P, P, P, P, S, S], I),
    "vc_float_rescore_device": ([S, S, P, P, P, P, P, S, S, P], I),
    "vc_landmarks_make": ([S, S, ctypes.c_uint64, P], I),
    "vc_landmarks_quantize": ([P, S, S, P, P], I),
    "vc_lm_index_create": ([S, S, S, ctypes.c_uint64, S, ctypes.POINTER(P)], I),
    "vc_lm_index_destroy": ([P], I),
    "vc_lm_index_add": ([P, P, S], I),
    "vc_lm_index_add_device": ([P, P, S, P], I),
    "vc_lm_index_prepare": ([P], I),
    "vc_lm_index_search": ([P, P, S, S, P, P, P, ctypes.POINTER(S)], I),
    "vc_lm_index_search_device": ([P, P, S, S, P, P, P, ctypes.POINTER(S), P], I),
    "vc_lm_index_query_rows": ([P, P, S, P, P], I),
    "vc_lm_index_export": ([P, P, P, P], I),
    "vc_lm_index_dim": ([P], S),
    "vc_lm_index_size": ([P], S),
    "vc_lm_index_scan_bits": ([P], S),
    "vc_lm_index_rerank_bits": ([P], S),
    "vc_lm_index_n_landmarks": ([P], S),
    "vc_lm_index_scan_dim": ([P], S),
    "vc_lm_index_is_prepared": ([P], I),
    "vc_lm_index_codes_bytes": ([P], S),
    "vc_lm_index_set_planes": ([P, I], I),
    "vc_lm_index_planes_active": ([P], I),
    "vc_lm_index_index_bytes": ([P], S),
    "vc_bucket_params_default": ([ctypes.POINTER(BucketParams)], None),
    "vc_bucketed_create": ([S, S, ctypes.POINTER(BucketParams), ctypes.POINTER(P)], I),
    "vc_bucketed_destroy": ([P], I),
    "vc_bucketed_calibrate": ([P, P, S], I),
    "vc_bucketed_calibrate_device": ([P, P, S, P], I),
    "vc_bucketed_add": ([P, P, S], I),
    "vc_bucketed_add_device": ([P, P, S, P], I),
    "vc_bucketed_prepare": ([P], I),
    "vc_bucketed_search": ([P, P, S, S, P, P, ctypes.POINTER(S)], I),
    "vc_bucketed_search_device": ([P, P, S, S, S, P, P, P, P, ctypes.POINTER(S), P], I),
    "vc_bucketed_route_device": ([P, P, S, S, P, P, P], I),
    "vc_bucketed_search_exhaustive_device": ([P, P, S, S, P, P, ctypes.POINTER(S), P], I),
    "vc_bucketed_dim": ([P], S),
    "vc_bucketed_size": ([P], S),
    "vc_bucketed_bits": ([P], S),
    "vc_bucketed_n_buckets": ([P], S),
    "vc_bucketed_is_prepared": ([P], I),
    "vc_bucketed_is_calibrated": ([P], I),
    "vc_bucketed_export": ([P, P, P, P, P, P, P, P], I),
    "vc_probe_param_count": ([], S),
    "vc_bucketed_probe_features": ([P, P, S, P, P], I),
    "vc_bucketed_probe_init": ([P, ctypes.c_uint64], I),
    "vc_bucketed_probe_set_weights": ([P, P], I),
    "vc_bucketed_probe_get_weights": ([P, P], I),
    "vc_bucketed_probe_train": ([P, P, P, S, P, S, F, P, P], I),
    "vc_bucketed_probe_grads": ([P, P, P, S, P, P], I),
    "vc_bucketed_probe_predict": ([P, P, S, P, P], I),
    "vc_bucketed_probe_enable": ([P, I, F], I),
    "vc_bucketed_probe_set_table": ([P, P, S, F, F], I),
    "vc_set_scan_tile": ([S], I),
    "vc_scan_tile_for": ([S, S, S], S),
    "vc_set_profiling": ([I], I),
    "vc_set_debug": ([ctypes.c_uint], I),
    "vc_last_scan_stats": ([ctypes.POINTER(ctypes.c_double), ctypes.POINTER(S), ctypes.POINTER(S)], I),
    "vc_npy_open": ([ctypes.c_char_p, ctypes.POINTER(P), ctypes.POINTER(S), ctypes.POINTER(S)], I),
    "vc_npy_data": ([P], ctypes.POINTER(ctypes.c_float)),
    "vc_npy_read_rows": ([P, S, S, P, ctypes.POINTER(S)], I),
    "vc_npy_close": ([P], I),
    "vc_npy_write_f32": ([ctypes.c_char_p, S, S, P], I),
    "vc_vcex_save": ([ctypes.c_char_p, ctypes.POINTER(VcexKey), P], I),
    "vc_vcex_load": ([ctypes.c_char_p, ctypes.POINTER(VcexKey), P, ctypes.POINTER(I)], I),
    "vc_vcex_source_label": ([ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, S, ctypes.POINTER(S)], I),
    "vc_vcex_cache_path": ([ctypes.POINTER(VcexKey), ctypes.c_char_p, S, ctypes.POINTER(S)], I),
}

for _name, (_args, _res) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.argtypes = _args
    _f.restype = _res

EXPORTED = tuple(_SIGS)


def check(status: int) -> None:
    """Raise the Python counterpart of the reference's exception for a status."""
    if status == VC_OK:
        return
    msg = (lib.vc_last_error() or b"").decode()
    if status == VC_ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    if status == VC_ERR_STATE:
        raise IndexStateError(msg)
    if status == VC_ERR_OUT_OF_RANGE:
        raise IndexError(msg)
    if status == VC_ERR_OOM:
        raise DeviceOutOfMemory(msg)
    raise VectorCacheError(msg)


def version() -> str:
    return lib.vc_version().decode()
