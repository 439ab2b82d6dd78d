"""Loads libhnsw_amd.so (the C ABI of include/hnsw_amd.h) and declares it.

The library is built in-tree by `make -C hnsw-index_amd` (or
__graft_entry__.build()).  There is no fallback: if the library is missing,
importing the package fails with an ImportError that says how to build it.
"""
import ctypes
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "libhnsw_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is not built; run `make -C {PKG_ROOT}` (hipcc, gfx950) first")

lib = ctypes.CDLL(LIB_PATH)

P = ctypes.c_void_p
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64
I32 = ctypes.c_int32
INT = ctypes.c_int


class Params(ctypes.Structure):
    _fields_ = [("m", U32), ("m0", U32), ("ef_construction", U32), ("selection", U32),
                ("extend_candidates", U32), ("keep_pruned", U32),
                ("level_multiplier", ctypes.c_double), ("seed", U64)]


class GraphDesc(ctypes.Structure):
    _fields_ = [("dim", U32), ("metric", U32), ("m", U32), ("m0", U32), ("n", U64),
                ("entry_id", U32), ("top_level", I32), ("vectors", P), ("keys", P),
                ("levels", P), ("layer0", P), ("upper_off", P), ("upper", P), ("deleted", P),
                ("codes", P), ("sq_offset", P), ("sq_scale", ctypes.c_float), ("live", U64)]


class FormatInfo(ctypes.Structure):
    _fields_ = [("kind", U32), ("version", U32), ("stored", U32), ("computed", U32),
                ("required", U64), ("limit", U64), ("what", ctypes.c_char * 48)]


class SearchStats(ctypes.Structure):
    _fields_ = [("queries", U64), ("n_dist", U64), ("n_upper", U64), ("n_expand", U64),
                ("n_rerank", U64), ("n_overflow", U64), ("n_visited_spill", U64),
                ("n_cand_spill", U64), ("n_redo", U64)]


class QuerySpec(ctypes.Structure):
    _fields_ = [("k", U32), ("ef", U32), ("precision", U32), ("filter", U32)]


class ServeConfig(ctypes.Structure):
    _fields_ = [("max_batch", U32), ("max_inflight", U32), ("k_max", U32), ("max_filters", U32),
                ("max_queue", U64)]


def _sig(name, restype, *args):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = list(args)


_sig("hh_last_error", ctypes.c_char_p)
_sig("hh_device_count", INT)
_sig("hh_abi_version", U32)
_sig("hh_params_default", None, U32, ctypes.POINTER(Params))
_sig("hh_graph_new", INT, U32, U32, ctypes.POINTER(Params), ctypes.POINTER(P))
_sig("hh_graph_free", None, P)
_sig("hh_graph_insert_batch", INT, P, P, P, U64, INT, ctypes.POINTER(U64))
_sig("hh_graph_remove", INT, P, U64, ctypes.POINTER(INT))
_sig("hh_graph_enable_quantization", INT, P)
_sig("hh_graph_set_quantizer", INT, P, P, ctypes.c_float, P)
_sig("hh_graph_disable_quantization", INT, P)
_sig("hh_graph_len", U64, P)
_sig("hh_graph_node_count", U64, P)
_sig("hh_graph_view", INT, P, ctypes.POINTER(GraphDesc))
_sig("hh_index_upload", INT, ctypes.POINTER(GraphDesc), INT, ctypes.POINTER(P))
_sig("hh_index_free", None, P)
_sig("hh_index_update_deleted", INT, P, P, U64)
_sig("hh_index_enable_bf16", INT, P)
_sig("hh_index_quantize", INT, P, P, ctypes.POINTER(ctypes.c_float))
_sig("hh_index_set_query_order", INT, P, INT)
_sig("hh_index_set_beam_width", INT, P, ctypes.c_uint32)
_sig("hh_index_beam_steps", INT, P, P)
_sig("hh_index_sync", INT, P, ctypes.POINTER(GraphDesc), P, U64)
_sig("hh_graph_take_dirty", INT, P, P, U64, ctypes.POINTER(U64))
_sig("hh_graph_append", INT, P, P, P, U64, ctypes.POINTER(U64), ctypes.POINTER(U64))
_sig("hh_graph_link_batch", INT, P, U64, U64, P, P, P, U32, INT)
_sig("hh_index_build_candidates", INT, P, U64, U64, U32, P, P, P)
_sig("hh_index_download_codes", INT, P, P)
_sig("hh_index_quantize_perdim", INT, P, P, P)
_sig("hh_index_download_perdim", INT, P, P, P)
_sig("hh_index_download_bf16", INT, P, P)
_sig("hh_index_stats", INT, P, ctypes.POINTER(SearchStats))
_sig("hh_index_reset_stats", INT, P)
_sig("hh_index_set_timing", INT, P, INT)
_sig("hh_index_autotune", INT, P, P, U64, U32, U32, U32, P, P, P, P, P, ctypes.POINTER(U32),
     ctypes.POINTER(INT))
_sig("hh_index_last_kernel_ms", INT, P, ctypes.POINTER(ctypes.c_float))
_sig("hh_index_gather_probe", INT, P, U32, ctypes.POINTER(ctypes.c_double),
     ctypes.POINTER(ctypes.c_float))
_sig("hh_search_batch_device", INT, P, P, U64, U32, U32, U32, P, P, P, P, P, P)
_sig("hh_search_batch_device_noalloc", INT, P, P, U64, U32, U32, U32, P, P, P, P, P, P)
_sig("hh_search_batch_spec_device", INT, P, P, U64, P, U32, U32, P, U32, P, P, P, P, P)
_sig("hh_search_batch_spec", INT, P, P, U64, P, U32, P, U32, P, P, P, P)
_sig("hh_search_plan_create", INT, P, P, U64, U32, U32, U32, P, P, U32, P, U32, P, P, P, P,
     ctypes.POINTER(P))
_sig("hh_search_plan_launch", INT, P, P)
_sig("hh_search_plan_free", None, P)
_sig("hh_serve_open", INT, P, ctypes.POINTER(ServeConfig), ctypes.POINTER(P))
_sig("hh_serve_register_filter", INT, P, P, ctypes.POINTER(U32))
_sig("hh_serve_submit", INT, P, P, ctypes.POINTER(QuerySpec), ctypes.POINTER(U64))
_sig("hh_serve_poll", INT, P, U64, INT, P, P, ctypes.POINTER(U32), ctypes.POINTER(I32),
     ctypes.POINTER(ctypes.c_double), ctypes.POINTER(INT))
_sig("hh_serve_stats", INT, P, ctypes.POINTER(U64), ctypes.POINTER(U64), ctypes.POINTER(U64))
_sig("hh_serve_close", INT, P)
_sig("hh_serve_replay", INT, P, P, P, U64, P, P, P, P, P, P)
_sig("hh_search_batch", INT, P, P, U64, U32, U32, U32, P, P, P, P, ctypes.POINTER(U64))
_sig("hh_brute_upload", INT, U32, U32, P, P, U64, INT, ctypes.POINTER(P), ctypes.POINTER(U64))
_sig("hh_brute_free", None, P)
_sig("hh_brute_len", U64, P)
_sig("hh_graph_info", I32, P, ctypes.POINTER(U32), ctypes.POINTER(U32), P)
_sig("hh_graph_memory_usage", U64, P)
_sig("hh_graph_to_bytes", I32, P, P, U64, ctypes.POINTER(U64))
_sig("hh_graph_from_bytes", I32, P, U64, U64, ctypes.POINTER(P), P)
_sig("hh_crc32", U32, P, U64)
_sig("hh_index_set_launch_shape", I32, P, U32, U32, INT)
_sig("hh_device_alloc", I32, INT, U64, ctypes.POINTER(P))
_sig("hh_device_free", I32, P)
_sig("hh_memcpy", I32, P, P, U64, INT, P)
_sig("hh_memset", I32, P, INT, U64)
_sig("hh_stream_create", I32, INT, ctypes.POINTER(P))
_sig("hh_stream_destroy", I32, P)
_sig("hh_stream_synchronize", I32, P)
_sig("hh_device_synchronize", I32, INT)
_sig("hh_brute_set_mfma", I32, P, I32)
_sig("hh_brute_mfma_stats", I32, P, ctypes.POINTER(U64), ctypes.POINTER(U64))
_sig("hh_brute_search_batch_device", INT, P, P, U64, U32, P, P, P, P, P, P)
_sig("hh_brute_search_batch_f64_device", INT, P, P, U64, U32, P, P, P, P, P)
_sig("hh_brute_search_batch", INT, P, P, U64, U32, P, P, P, P, ctypes.POINTER(U64))
_sig("hh_brute_search_batch_f64", INT, P, P, U64, U32, P, P, P, ctypes.POINTER(U64))
_sig("hh_synthetic", INT, U32, U64, U64, U32, U32, ctypes.c_double, ctypes.c_double, U64, INT,
     P, P, P)
_sig("hh_cpu_distance", ctypes.c_float, U32, P, P, U64)
_sig("hh_recall_at_k", ctypes.c_double, U64, P, P, U32, P, P, U32, U32)


def last_error():
    return lib.hh_last_error().decode("utf-8", "replace")
