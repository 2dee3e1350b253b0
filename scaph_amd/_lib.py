"""ctypes bindings of libscaph_hip.so (include/scaph_hip.h) and
libscaph_host.so (include/scaph_engine.h).

The libraries are built in-tree by ``__graft_entry__.build()`` (make -C
scaph_amd).  Loading fails loudly when they are missing: there is no
fallback path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIB = os.path.join(HERE, "libscaph_hip.so")
HOST_LIB = os.path.join(HERE, "libscaph_host.so")

UNREACHED = 0xFFFFFFFF
STATUS_NOT_READY = 600


class LibraryMissing(RuntimeError):
    pass


class ScaphError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"status {status}: {msg}")
        self.status = status


class Page(ctypes.Structure):
    """src/page.h:22-29 Page<uint32>."""
    _fields_ = [("left", ctypes.c_uint32), ("right", ctypes.c_uint32),
                ("nodenum", ctypes.c_uint32), ("edgenum", ctypes.c_uint32)]


class GraphDesc(ctypes.Structure):
    _fields_ = [("page_ptr", ctypes.c_void_p), ("pages", ctypes.c_void_p),
                ("num_pages", ctypes.c_uint32), ("num_vertices", ctypes.c_uint32)]


class LevelCounters(ctypes.Structure):
    _fields_ = [("n_next", ctypes.c_ulonglong), ("m_next", ctypes.c_ulonglong),
                ("q_len", ctypes.c_ulonglong), ("scanned", ctypes.c_ulonglong),
                ("examined", ctypes.c_ulonglong), ("reserved", ctypes.c_ulonglong)]


class DeviceProps(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 256), ("arch", ctypes.c_char * 64),
                ("compute_units", ctypes.c_int), ("xcds", ctypes.c_int),
                ("wave_size", ctypes.c_int), ("lds_per_block", ctypes.c_size_t),
                ("total_mem", ctypes.c_size_t), ("l2_bytes", ctypes.c_size_t),
                ("clock_khz", ctypes.c_int), ("mem_clock_khz", ctypes.c_int),
                ("mem_bus_width", ctypes.c_int)]


class ExploreRecord(ctypes.Structure):
    _fields_ = [("n_f", ctypes.c_uint64), ("m_f", ctypes.c_uint64), ("n_next", ctypes.c_uint64),
                ("level", ctypes.c_uint32), ("t_push_ms", ctypes.c_float),
                ("t_pull_ms", ctypes.c_float), ("pages_active_next", ctypes.c_float),
                ("x", ctypes.c_float * 16)]


class MsTune(ctypes.Structure):
    _fields_ = [("push_div", ctypes.c_uint32), ("push_mdiv", ctypes.c_uint32),
                ("pull_block", ctypes.c_uint32), ("use_graph", ctypes.c_uint32),
                ("segment", ctypes.c_uint32), ("pull_grid", ctypes.c_uint32),
                ("persistent", ctypes.c_uint32)]


class MsStats(ctypes.Structure):
    _fields_ = [("ms", ctypes.c_double), ("depth", ctypes.c_uint32), ("queries", ctypes.c_uint32),
                ("pairs", ctypes.c_uint64)]


class MsCounters(ctypes.Structure):
    _fields_ = [("n_active", ctypes.c_uint64), ("bits", ctypes.c_uint64),
                ("m_active", ctypes.c_uint64), ("q_len", ctypes.c_uint64),
                ("heavy_len", ctypes.c_uint64), ("overflow", ctypes.c_uint64),
                ("pad", ctypes.c_uint64 * 2)]


class BatchEntry(ctypes.Structure):
    _fields_ = [("vbase", ctypes.c_uint64), ("ebase", ctypes.c_uint64), ("seed", ctypes.c_uint64),
                ("n", ctypes.c_uint32), ("scale", ctypes.c_uint32), ("src", ctypes.c_uint32),
                ("edge_factor", ctypes.c_uint32)]


class BatchTune(ctypes.Structure):
    _fields_ = [("block", ctypes.c_uint32), ("wgs_per_cu", ctypes.c_uint32),
                ("use_graph", ctypes.c_uint32)]


class BfsStats(ctypes.Structure):
    _fields_ = [("src", ctypes.c_uint32), ("depth", ctypes.c_uint32), ("ms", ctypes.c_double),
                ("n_reached", ctypes.c_uint64), ("m_reached", ctypes.c_uint64),
                ("page_execs", ctypes.c_uint64), ("page_execs_cached", ctypes.c_uint64),
                ("pages_full", ctypes.c_uint64), ("pages_compacted", ctypes.c_uint64),
                ("bytes_h2d", ctypes.c_uint64), ("h2d_ms", ctypes.c_double),
                ("pages_prefetched", ctypes.c_uint64), ("pages_direct", ctypes.c_uint64)]


class LevelRecord(ctypes.Structure):
    _fields_ = [("level", ctypes.c_uint32), ("dir", ctypes.c_int32),
                ("n_f", ctypes.c_uint64), ("m_f", ctypes.c_uint64), ("m_u", ctypes.c_uint64),
                ("n_next", ctypes.c_uint64), ("m_next", ctypes.c_uint64),
                ("scanned", ctypes.c_uint64), ("examined", ctypes.c_uint64),
                ("ms", ctypes.c_float), ("kernel_ms", ctypes.c_float),
                ("page_hist", ctypes.c_float * 12),
                ("pages_active", ctypes.c_uint32), ("pages_cached", ctypes.c_uint32),
                ("pages_full", ctypes.c_uint32), ("pages_compacted", ctypes.c_uint32),
                ("pages_prefetched", ctypes.c_uint32), ("pages_direct", ctypes.c_uint32),
                ("bytes_h2d", ctypes.c_uint64), ("h2d_ms", ctypes.c_float),
                ("logit", ctypes.c_float), ("sparse", ctypes.c_uint32), ("pad", ctypes.c_uint32)]


_hip = None
_host = None

P = ctypes.c_void_p
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64
I32 = ctypes.c_int
SZ = ctypes.c_size_t
DBL = ctypes.c_double
CHARP = ctypes.c_char_p


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


def hip():
    """The kernel library (loads it on first use)."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB):
            raise LibraryMissing(f"{HIP_LIB} is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(HIP_LIB, mode=ctypes.RTLD_GLOBAL)
        _sig(lib, "scaph_error_string", CHARP, [I32])
        _sig(lib, "scaph_init", I32, [I32])
        _sig(lib, "scaph_device_count", I32, [ctypes.POINTER(I32)])
        _sig(lib, "scaph_get_device_props", I32, [I32, ctypes.POINTER(DeviceProps)])
        for n in ("scaph_stream_destroy", "scaph_stream_query", "scaph_stream_sync",
                  "scaph_event_destroy", "scaph_event_query", "scaph_event_sync",
                  "scaph_dev_free", "scaph_host_pinned_free", "scaph_graph_destroy",
                  "scaph_graph_capture_begin"):
            _sig(lib, n, I32, [P])
        _sig(lib, "scaph_stream_create", I32, [ctypes.POINTER(P)])
        _sig(lib, "scaph_stream_create_cu_limited", I32, [ctypes.POINTER(P), U32])
        _sig(lib, "scaph_hw_ids", I32, [P, U32, P])
        _sig(lib, "scaph_event_create", I32, [ctypes.POINTER(P)])
        _sig(lib, "scaph_event_record", I32, [P, P])
        _sig(lib, "scaph_stream_wait_event", I32, [P, P])
        _sig(lib, "scaph_event_elapsed_ms", I32, [P, P, ctypes.POINTER(ctypes.c_float)])
        _sig(lib, "scaph_device_sync", I32, [])
        _sig(lib, "scaph_mem_info", I32, [ctypes.POINTER(SZ), ctypes.POINTER(SZ)])
        _sig(lib, "scaph_dev_alloc", I32, [ctypes.POINTER(P), SZ])
        _sig(lib, "scaph_host_pinned_alloc", I32, [ctypes.POINTER(P), SZ])
        for n in ("scaph_copy_h2d_async", "scaph_copy_d2h_async", "scaph_copy_d2d_async"):
            _sig(lib, n, I32, [P, P, SZ, P])
        _sig(lib, "scaph_fill_u32", I32, [P, U32, SZ, P])
        _sig(lib, "scaph_checksum_u32", I32, [P, SZ, P, P])
        _sig(lib, "scaph_bfs_init", I32, [P, P, P, U32, U32, P, P, P])
        _sig(lib, "scaph_bitmap_from_zero_degree", I32, [P, P, U32, P])
        _sig(lib, "scaph_push_workspace_bytes", SZ, [U32, ctypes.c_ulonglong])
        _sig(lib, "scaph_bfs_push", I32,
             [ctypes.POINTER(GraphDesc), P, P, P, P, U32, ctypes.c_ulonglong, P, P, U32, P, SZ, P,
              P])
        _sig(lib, "scaph_bfs_push_noalloc_max_bytes", SZ, [])
        _sig(lib, "scaph_bfs_push_noalloc", I32,
             [ctypes.POINTER(GraphDesc), P, P, P, P, U32, ctypes.c_ulonglong, P, P, U32, P, P])
        _sig(lib, "scaph_pull_workspace_bytes", SZ, [ctypes.c_ulonglong])
        _sig(lib, "scaph_bfs_pull", I32, [ctypes.POINTER(GraphDesc), P, P, U32, P, P, P, P, P, P,
                                          U32, ctypes.c_ulonglong, P, SZ, P, P])
        _sig(lib, "scaph_list_heads_from_csr", I32, [P, P, U32, P, P, P])
        _sig(lib, "scaph_live_window_base", I32, [P, U32, P, P])
        _sig(lib, "scaph_list_heads_live", I32, [P, P, U32, P, P, P, P, P])
        _sig(lib, "scaph_list_heads_indexed", I32, [P, P, U32, P, P, U32, P, P, P, P])
        _sig(lib, "scaph_frontier_queue_to_bitmap", I32, [P, U32, P, U32, P])
        _sig(lib, "scaph_frontier_bitmap_to_queue", I32, [P, U32, P, P, P])
        _sig(lib, "scaph_page_counts", I32, [P, U32, P, P, I32, P, P, P, P])
        _sig(lib, "scaph_page_unvisited", I32, [P, U32, P, P, P, P])
        _sig(lib, "scaph_bfs_push_page", I32,
             [P, U32, ctypes.c_ulonglong, U32, P, P, P, P, P, P, U32, P, SZ, P, P, P])
        _sig(lib, "scaph_xcc_ids", I32, [P, U32, P])
        _sig(lib, "scaph_bfs_validate", I32,
             [ctypes.POINTER(GraphDesc), ctypes.POINTER(GraphDesc), P, P, U32, P, P])
        _sig(lib, "scaph_gen_rmat_keys", I32, [P, U64, U64, U32, U32, U32, U32, U64, P])
        _sig(lib, "scaph_gen_er_keys", I32, [P, U64, U64, U32, U64, P])
        _sig(lib, "scaph_csr_from_keys", I32,
             [P, P, U64, U32, I32, P, P, P, P, ctypes.POINTER(U64), P])
        _sig(lib, "scaph_pages_from_csr", I32, [P, P, P, P, U32, U32, P])
        _sig(lib, "scaph_mlp_forward", I32, [P, P, U32, P, P, P])
        _sig(lib, "scaph_mlp_backward", I32, [P, P, P, P, P, U32, ctypes.c_float, P, P, P])
        _sig(lib, "scaph_mlp_sgd", I32, [P, P, P, U32, ctypes.c_float, ctypes.c_float, P, P])
        _sig(lib, "scaph_frontier_page_counts", I32, [P, P, P, P, U32, U32, P, P])
        _sig(lib, "scaph_mlp_decide", I32, [P, P, P, P, U32, P, ctypes.c_ulonglong, U32, P, P])
        _sig(lib, "scaph_mlp_decide_publish", I32,
             [P, P, P, U32, ctypes.c_ulonglong, U32, U32, U32, P, P, U32, P])
        _sig(lib, "scaph_host_mapped_alloc", I32,
             [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p), SZ])
        _sig(lib, "scaph_publish_words", I32, [P, U32, U32, P, P, U32, P])
        _sig(lib, "scaph_mlp_decide_frontier", I32,
             [P, P, P, P, P, U32, U32, P, ctypes.c_ulonglong, U32, P, P, P])
        _sig(lib, "scaph_graph_capture_end", I32, [P, ctypes.POINTER(P)])
        _sig(lib, "scaph_graph_launch", I32, [P, P])
        _hip = lib
    return _hip


def host():
    """The host engine library (loads the kernel library first)."""
    global _host
    if _host is None:
        hip()
        if not os.path.exists(HOST_LIB):
            raise LibraryMissing(f"{HOST_LIB} is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(HOST_LIB, mode=ctypes.RTLD_GLOBAL)
        E = P
        _sig(lib, "scaph_engine_create", I32,
             [I32, ctypes.POINTER(CHARP), ctypes.POINTER(E), CHARP, SZ])
        _sig(lib, "scaph_engine_destroy", None, [E])
        _sig(lib, "scaph_engine_load", I32, [E, CHARP, SZ])
        _sig(lib, "scaph_engine_load_csr", I32, [E, U32, P, P, I32, CHARP, SZ])
        _sig(lib, "scaph_engine_gen_rmat_device", I32,
             [E, U32, U32, DBL, DBL, DBL, U64, I32, CHARP, SZ])
        _sig(lib, "scaph_engine_gen_er_device", I32, [E, U32, U64, U64, I32, CHARP, SZ])
        _sig(lib, "scaph_engine_gen_zipf", I32, [E, U32, DBL, DBL, U32, U64, CHARP, SZ])
        _sig(lib, "scaph_engine_info", I32,
             [E, ctypes.POINTER(U32), ctypes.POINTER(U64), ctypes.POINTER(U32),
              ctypes.POINTER(I32), ctypes.POINTER(I32)])
        _sig(lib, "scaph_engine_bfs", I32, [E, U32, ctypes.POINTER(BfsStats), CHARP, SZ])
        _sig(lib, "scaph_engine_wcc", I32, [E, ctypes.POINTER(BfsStats), CHARP, SZ])
        _sig(lib, "scaph_engine_sssp", I32, [E, U32, ctypes.POINTER(BfsStats), CHARP, SZ])
        _sig(lib, "scaph_engine_set_weights", I32, [E, P, U64, CHARP, SZ])
        _sig(lib, "scaph_engine_set_schedule", I32, [E, CHARP, DBL, DBL, CHARP, SZ])
        _sig(lib, "scaph_engine_set_page_levels", I32, [E, P, U32, CHARP, SZ])
        _sig(lib, "scaph_engine_set_model", I32, [E, P, P, P, ctypes.c_float, CHARP, SZ])
        _sig(lib, "scaph_engine_level_records", I32,
             [E, ctypes.POINTER(LevelRecord), U32, ctypes.POINTER(U32)])
        _sig(lib, "scaph_engine_levels", I32, [E, P])
        _sig(lib, "scaph_engine_explore", I32,
             [E, U32, I32, ctypes.POINTER(ExploreRecord), U32, ctypes.POINTER(U32), CHARP, SZ])
        _sig(lib, "scaph_engine_validate", I32, [E, U32, P])
        _sig(lib, "scaph_engine_msbfs", I32,
             [E, P, U32, ctypes.POINTER(MsTune), ctypes.POINTER(MsStats), CHARP, SZ])
        _sig(lib, "scaph_engine_msbfs_levels", I32, [E, U32, P, CHARP, SZ])
        _sig(lib, "scaph_engine_msbfs_counters", I32,
             [E, ctypes.POINTER(MsCounters), U32, ctypes.POINTER(U32)])
        _sig(lib, "scaph_engine_msbfs_autotune", I32,
             [E, P, U32, ctypes.POINTER(MsTune), ctypes.POINTER(DBL), ctypes.POINTER(MsTune),
              ctypes.POINTER(DBL), U32, ctypes.POINTER(U32), CHARP, SZ])
        _sig(lib, "scaph_engine_out_degrees", I32, [E, P])
        _sig(lib, "scaph_engine_host_csr", I32, [E, P, P])
        _sig(lib, "scaph_engine_check", I32, [E, U32, ctypes.POINTER(U64), CHARP, SZ])
        _sig(lib, "scaph_engine_graph_desc", I32, [E, I32, ctypes.POINTER(GraphDesc)])
        _sig(lib, "scaph_engine_device_degree", I32, [E, ctypes.POINTER(P)])
        _sig(lib, "scaph_batch_create_rmat", I32,
             [U32, U32, U32, U32, U32, U64, I32, ctypes.POINTER(P), CHARP, SZ])
        _sig(lib, "scaph_batch_destroy", None, [P])
        _sig(lib, "scaph_batch_info", I32,
             [P, ctypes.POINTER(U32), ctypes.POINTER(U64), ctypes.POINTER(U64)])
        _sig(lib, "scaph_batch_graph_info", I32, [P, U32, ctypes.POINTER(BatchEntry)])
        _sig(lib, "scaph_batch_run", I32,
             [P, ctypes.POINTER(BatchTune), ctypes.POINTER(DBL), CHARP, SZ])
        _sig(lib, "scaph_batch_levels", I32, [P, U32, P])
        _sig(lib, "scaph_batch_csr", I32, [P, U32, P, P, U64, ctypes.POINTER(U64)])
        _sig(lib, "scaph_batch_autotune", I32,
             [P, ctypes.POINTER(BatchTune), ctypes.POINTER(DBL), ctypes.POINTER(BatchTune),
              ctypes.POINTER(DBL), U32, ctypes.POINTER(U32), CHARP, SZ])
        _sig(lib, "scaph_host_page_cuts", I32, [P, U32, U64, I32, P, U32, ctypes.POINTER(U32)])
        _sig(lib, "scaph_host_gen_rmat", I32,
             [U32, U32, DBL, DBL, DBL, U64, P, P, U64, ctypes.POINTER(U64)])
        _sig(lib, "scaph_host_gen_er", I32, [U32, U64, U64, P, P, U64, ctypes.POINTER(U64)])
        _sig(lib, "scaph_host_write_graph", I32, [CHARP, U32, P, P, U64, I32, I32, CHARP, SZ])
        _sig(lib, "scaph_host_cpu_bfs", I32, [U32, P, P, U32, U32, P])
        _sig(lib, "scaph_host_plan_pool", I32,
             [ctypes.c_int64, U32, U32, U64, ctypes.POINTER(ctypes.c_int64)])
        _sig(lib, "scaph_engine_pool", I32,
             [E, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(U64)])
        _host = lib
    return _host


def check(status, msg=""):
    if status != 0:
        text = msg
        if not text:
            try:
                text = hip().scaph_error_string(status).decode()
            except Exception:  # pragma: no cover
                text = "error"
        raise ScaphError(status, text)
