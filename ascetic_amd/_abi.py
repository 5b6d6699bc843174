"""ctypes declarations of the C ABI in include/ascetic.h.

The structures below mirror the header field by field; `load()` opens the
in-tree libascetic.so and fails loudly when it is missing.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libascetic.so")
CLI_PATH = os.path.join(_HERE, "ptgraph")

# status codes
OK = 0
ERR_INVALID = -1
ERR_IO = -2
ERR_FORMAT = -3
ERR_NOMEM = -4
ERR_DEVICE = -5
ERR_NOGPU = -6
ERR_RANGE = -7
ERR_UNSUPPORTED = -8
ERR_BUSY = -9
ERR_STATE = -10

FMT_AUTO, FMT_BCSR, FMT_BWCSR, FMT_BCSC, FMT_LEGACY32, FMT_LEGACY32W = range(6)
KIND_CSR, KIND_WCSR, KIND_CSC = range(3)
VALUES_REFERENCE, VALUES_ZERO_BASED = 0, 1
STREAM_GATHER, STREAM_ZEROCOPY, STREAM_HYBRID = 0, 1, 2
PR_NORMALISED, PR_REFERENCE = 0, 1
PR_AUTO, PR_PULL, PR_BLOCKED, PR_TILED = 0, 1, 2, 3
BFS_AUTO, BFS_TOP_DOWN = 0, 1
CC_AUTO, CC_LABEL_PROP = 0, 1
QUERY_BFS, QUERY_SSSP = 0, 1

u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)


class Config(C.Structure):
    _fields_ = [
        ("device", C.c_int),
        ("advise_rate", C.c_double),
        ("edge_budget_bytes", C.c_uint64),
        ("part_bytes", C.c_uint64),
        ("device_mem_limit", C.c_uint64),
        ("gather_threads", C.c_int),
        ("stream_mode", C.c_int),
        ("value_convention", C.c_int),
        ("heavy_threshold", C.c_int),
        ("pr_group", C.c_int),
        ("sssp_delta", C.c_int),
        ("verbose", C.c_int),
        ("profile", C.c_int),
        ("stream_pack", C.c_int),
        ("bfs_direction", C.c_int),
        ("cc_method", C.c_int),
        ("dense_threshold", C.c_double),
        ("stream_overlap", C.c_int),
        ("bfs_persistent", C.c_int),
        ("release_host_edges", C.c_int),
        ("part_buffers", C.c_int),
        ("pad_cfg", C.c_int),
        ("max_pinned_bytes", C.c_uint64),
    ]


class CsrView(C.Structure):
    _fields_ = [
        ("num_vertices", C.c_uint64),
        ("num_edges", C.c_uint64),
        ("kind", C.c_int),
        ("has_sentinel", C.c_int),
        ("off64", u64p),
        ("off32", u32p),
        ("edges", u32p),
        ("out_degree", u32p),
    ]


class GraphInfo(C.Structure):
    _fields_ = [
        ("num_vertices", C.c_uint64),
        ("num_edges", C.c_uint64),
        ("kind", C.c_int),
        ("offset_bytes", C.c_int),
        ("static_edges", C.c_uint64),
        ("static_vertices", C.c_uint64),
        ("ondemand_edges", C.c_uint64),
        ("part_edges", C.c_uint64),
        ("device_bytes", C.c_uint64),
        ("max_degree", C.c_uint32),
        ("max_degree_vertex", C.c_uint32),
        ("symmetric", C.c_int),
        ("tail_pinned", C.c_int),
        ("host_bytes", C.c_uint64),
    ]


class KernelStats(C.Structure):
    _fields_ = [
        ("launches", C.c_uint32),
        ("pad", C.c_uint32),
        ("ms", C.c_double),
        ("edges", C.c_uint64),
        ("vertices", C.c_uint64),
    ]


KERNEL_NAMES = ("expand_local", "expand_heavy", "expand_part", "pr_pull", "bottom_up", "union_find", "bfs_persistent",
                "expand_part_host")
PATH_NAMES = ("frontier", "direction_opt", "union_find")


class Stats(C.Structure):
    _fields_ = [
        ("total_ms", C.c_double),
        ("pre_ms", C.c_double),
        ("static_ms", C.c_double),
        ("overload_ms", C.c_double),
        ("overload_move_ms", C.c_double),
        ("fill_ms", C.c_double),
        ("iterations", C.c_uint32),
        ("parts", C.c_uint32),
        ("node_sum", C.c_uint64),
        ("edges_traversed", C.c_uint64),
        ("static_edges_traversed", C.c_uint64),
        ("streamed_edges", C.c_uint64),
        ("streamed_bytes", C.c_uint64),
        ("dense_edges", C.c_uint64),
        ("vertices_processed", C.c_uint64),
        ("reached", C.c_uint64),
        ("reached_degree_sum", C.c_uint64),
        ("kernel_ms", C.c_double),
        ("pr_residual", C.c_double),
        ("host_cores", C.c_int),
        ("path", C.c_int),
        ("kern", KernelStats * 8),
        ("reused_edges", C.c_uint64),
    ]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_ if name != "kern"}
        d["path"] = PATH_NAMES[self.path] if 0 <= self.path < len(PATH_NAMES) else self.path
        d["kernels"] = {KERNEL_NAMES[i]: {n: getattr(self.kern[i], n) for n in ("launches", "ms", "edges", "vertices")}
                        for i in range(len(KERNEL_NAMES))}
        return d


class PrOpts(C.Structure):
    _fields_ = [
        ("mode", C.c_int),
        ("damping", C.c_double),
        ("iterations", C.c_int),
        ("threshold", C.c_double),
        ("max_iterations", C.c_int),
        ("method", C.c_int),
        ("tile_src_bits", C.c_int),
        ("tile_dst_bits", C.c_int),
    ]


class Plan(C.Structure):
    _fields_ = [
        ("total_device_edges", C.c_uint64),
        ("advise_k", C.c_double),
        ("static_edges", C.c_uint64),
        ("static_vertices", C.c_uint64),
        ("ondemand_edges", C.c_uint64),
        ("part_edges", C.c_uint64),
    ]


class GenParams(C.Structure):
    _fields_ = [
        ("type", C.c_int),
        ("scale", C.c_int),
        ("num_vertices", C.c_uint64),
        ("num_samples", C.c_uint64),
        ("edge_factor", C.c_int),
        ("a", C.c_double),
        ("b", C.c_double),
        ("c", C.c_double),
        ("seed", C.c_uint64),
        ("symmetrize", C.c_int),
        ("weighted", C.c_int),
        ("wmin", C.c_uint32),
        ("wmax", C.c_uint32),
        ("offset_bytes", C.c_int),
        ("csc", C.c_int),
        ("threads", C.c_int),
    ]


class EdgeListOpts(C.Structure):
    _fields_ = [
        ("num_vertices", C.c_uint64),
        ("symmetrize", C.c_int),
        ("dedup", C.c_int),
        ("keep_self_loops", C.c_int),
        ("one_based", C.c_int),
        ("weighted", C.c_int),
        ("csc", C.c_int),
        ("wmin", C.c_uint32),
        ("wmax", C.c_uint32),
        ("seed", C.c_uint64),
        ("threads", C.c_int),
        ("pad", C.c_int),
    ]


class HostCsr(C.Structure):
    _fields_ = [
        ("num_vertices", C.c_uint64),
        ("num_edges", C.c_uint64),
        ("kind", C.c_int),
        ("off", u64p),
        ("edges", u32p),
        ("out_degree", u32p),
    ]


class ServerOpts(C.Structure):
    _fields_ = [
        ("lanes", C.c_int),
        ("state_budget_bytes", C.c_uint64),
        ("queue_capacity", C.c_int),
        ("bfs_direction", C.c_int),
        ("step_graphs", C.c_int),
        ("pad", C.c_int),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64),
        ("completed", C.c_uint64),
        ("steps", C.c_uint64),
        ("lanes_bfs", C.c_int),
        ("lanes_sssp", C.c_int),
        ("busy_ms", C.c_double),
        ("bottom_up_steps", C.c_uint64),
        ("fallbacks", C.c_uint64),
        ("graph_steps", C.c_uint64),
    ]


class QStats(C.Structure):
    _fields_ = [
        ("submit_to_admit_ms", C.c_double),
        ("submit_to_done_ms", C.c_double),
        ("iterations", C.c_uint32),
        ("lane", C.c_uint32),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("heavy_threshold", C.c_int),
        ("blocks_per_cu", C.c_int),
        ("pr_group", C.c_int),
        ("sssp_delta", C.c_int),
        ("serve_lanes", C.c_int),
        ("pr_tile_src_bits", C.c_int),
        ("score_ms", C.c_double),
        ("untuned_ms", C.c_double),
        ("fingerprint", C.c_uint64),
    ]


P = C.c_void_p
PP = C.POINTER(C.c_void_p)

# name -> (restype, argtypes); every function declared in include/ascetic.h
SIGNATURES = {
    "asc_strerror": (C.c_char_p, [C.c_int]),
    "asc_abi_version": (C.c_int, []),
    "asc_config_init": (None, [C.POINTER(Config)]),
    "asc_ctx_create": (C.c_int, [C.POINTER(Config), PP]),
    "asc_ctx_destroy": (C.c_int, [P]),
    "asc_device_count": (C.c_int, []),
    "asc_ctx_synchronize": (C.c_int, [P]),
    "asc_measure_link": (C.c_int, [P, C.c_uint64, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "asc_copy_probe": (C.c_int, [P, C.c_uint64, C.POINTER(C.c_double)]),
    "asc_graph_load": (C.c_int, [P, C.c_char_p, C.c_int, PP]),
    "asc_graph_from_csr": (C.c_int, [P, C.POINTER(CsrView), PP]),
    "asc_graph_destroy": (C.c_int, [P]),
    "asc_graph_get_info": (C.c_int, [P, C.POINTER(GraphInfo)]),
    "asc_write_graph": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(CsrView)]),
    "asc_graph_replan": (C.c_int, [P, C.c_uint64, C.c_double]),
    "asc_bfs": (C.c_int, [P, C.c_uint32, u32p, C.POINTER(Stats)]),
    "asc_sssp": (C.c_int, [P, C.c_uint32, u32p, C.POINTER(Stats)]),
    "asc_bfs_noalloc": (C.c_int, [P, C.c_uint32, C.c_void_p, C.c_void_p]),
    "asc_cc": (C.c_int, [P, u32p, C.POINTER(Stats)]),
    "asc_pr_opts_init": (None, [C.POINTER(PrOpts)]),
    "asc_pr": (C.c_int, [P, C.POINTER(PrOpts), C.POINTER(C.c_float), C.POINTER(C.c_double),
                         C.POINTER(C.c_double), C.POINTER(Stats)]),
    "asc_plan_partition": (C.c_int, [C.c_uint64, C.c_uint64, u64p, C.c_uint64, C.c_uint64, C.c_int,
                                     C.c_double, C.c_uint64, C.c_uint64, C.POINTER(Plan)]),
    "asc_split_parts": (C.c_int, [u64p, C.c_uint64, C.c_uint64, u64p, C.c_uint64, u64p]),
    "asc_gen_params_init": (None, [C.POINTER(GenParams)]),
    "asc_generate": (C.c_int, [C.POINTER(GenParams), C.POINTER(HostCsr)]),
    "asc_generate_device": (C.c_int, [C.c_void_p, C.POINTER(GenParams), C.POINTER(HostCsr)]),
    "asc_read_graph": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(HostCsr)]),
    "asc_edge_list_opts_init": (None, [C.POINTER(EdgeListOpts)]),
    "asc_read_edge_list": (C.c_int, [C.c_char_p, C.POINTER(EdgeListOpts), C.POINTER(HostCsr)]),
    "asc_host_csr_free": (None, [C.POINTER(HostCsr)]),
    "asc_host_csr_view": (None, [C.POINTER(HostCsr), C.POINTER(CsrView)]),
    "asc_graph_adopt": (C.c_int, [P, C.POINTER(HostCsr), C.c_int, PP]),
    "asc_graph_borrow_csr": (C.c_int, [P, C.POINTER(CsrView), PP]),
    "asc_graph_generate": (C.c_int, [P, C.POINTER(GenParams), PP]),
    "asc_server_opts_init": (None, [C.POINTER(ServerOpts)]),
    "asc_server_create": (C.c_int, [P, C.POINTER(ServerOpts), PP]),
    "asc_server_destroy": (C.c_int, [P]),
    "asc_submit": (C.c_int, [P, C.c_int, C.c_uint32, u32p, u64p]),
    "asc_wait": (C.c_int, [P, C.c_uint64, C.POINTER(QStats)]),
    "asc_server_drain": (C.c_int, [P]),
    "asc_server_get_stats": (C.c_int, [P, C.POINTER(ServerStats)]),
    "asc_autotune": (C.c_int, [P, C.c_double, C.POINTER(Tuning)]),
    "asc_tuning_save": (C.c_int, [C.c_char_p, P, C.POINTER(Tuning)]),
    "asc_tuning_load": (C.c_int, [C.c_char_p, P, C.POINTER(Tuning)]),
    "asc_tuning_apply": (C.c_int, [P, C.POINTER(Tuning)]),
}

# diagnostics exported beside the header's functions
EXTRA = {
    "asc_last_error": (C.c_char_p, []),
}

_lib = None


def load():
    """Opens libascetic.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in {**SIGNATURES, **EXTRA}.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class AscError(RuntimeError):
    def __init__(self, code, what=""):
        lib = load()
        msg = lib.asc_strerror(code).decode()
        detail = lib.asc_last_error().decode()
        super().__init__(f"{what}: {msg} ({code}) {detail}".strip())
        self.code = code


def check(code, what=""):
    if code != OK:
        raise AscError(code, what)
    return code
