"""ORACLE -- test infrastructure, not product code.

ctypes wrapper of liboracle.so (oracle.c): the CPU restatement of the
reference's hot path and the textbook algorithms that pin it.  Imported only
by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
f64p = C.POINTER(C.c_double)


class Graph(C.Structure):
    _fields_ = [("V", C.c_uint64), ("E", C.c_uint64), ("off", u64p), ("dst", u32p), ("w", u32p),
                ("stride", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("node_sum", C.c_uint64), ("edges_relaxed", C.c_uint64),
                ("overload_edges", C.c_uint64), ("parts", C.c_uint32)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Plan(C.Structure):
    _fields_ = [("total_gpu_size", C.c_uint64), ("advise_k", C.c_double), ("max_partition_size", C.c_uint64),
                ("static_vertices", C.c_uint64), ("part_overload_size", C.c_uint64),
                ("overload_size", C.c_uint64)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "-B"], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        G = C.POINTER(Graph)
        S = C.POINTER(Stats)
        sig = {
            "orc_bfs_ref": [G, C.c_uint32, C.c_uint64, C.c_uint64, C.c_int, u32p, S],
            "orc_sssp_ref": [G, C.c_uint32, C.c_uint64, C.c_uint64, C.c_int, u32p, S],
            "orc_cc_ref": [G, C.c_uint64, C.c_uint64, C.c_int, u32p, S],
            "orc_pr_ref": [G, u32p, C.c_double, C.c_double, C.c_int, f64p, S],
            "orc_pr_norm": [G, u32p, C.c_double, C.c_int, f64p, f64p],
            "orc_bfs_queue": [G, C.c_uint32, u32p],
            "orc_dijkstra": [G, C.c_uint32, u32p],
            "orc_cc_minreach": [G, u32p],
            "orc_cc_unionfind": [G, u32p],
            "orc_plan_reference": [C.c_uint64, C.c_uint64, u64p, C.c_uint64, C.c_uint64, C.c_int, C.c_double,
                                   C.POINTER(Plan), C.POINTER(C.c_uint8)],
            "orc_split_reference": [u64p, u32p, C.c_uint32, C.c_uint64, C.c_uint64, u32p, u32p, u64p, C.c_int],
            "orc_set_fault": [C.c_int],
            "orc_set_iteration_limit": [C.c_int],
            "orc_set_edge_limit": [C.c_uint64],
            "orc_last_pulled": [],
            "orc_set_num_threads": [C.c_int],
            "orc_gather_ref": [G, u32p, C.c_uint64, C.c_int, u32p, f64p],
            "orc_rmat_pairs": [C.c_uint64, C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_uint32,
                               C.c_uint32, C.c_uint64, u32p, u32p, u32p],
        }
        void = ("orc_set_fault", "orc_set_iteration_limit", "orc_set_num_threads", "orc_set_edge_limit")
        for name, args in sig.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = (None if name in void else C.c_longlong if name == "orc_gather_ref"
                         else C.c_uint64 if name == "orc_last_pulled" else C.c_int)
        _lib = lib
    return _lib


class CSR:
    """Keeps numpy arrays alive behind an oracle Graph struct.

    off: V+1 offsets; dst: E ids; w: E weights or None."""

    def __init__(self, off, dst, w=None):
        self.off = np.ascontiguousarray(off, dtype=np.uint64)
        self.dst = np.ascontiguousarray(dst, dtype=np.uint32)
        self.w = None if w is None else np.ascontiguousarray(w, dtype=np.uint32)
        self.V = len(self.off) - 1
        self.E = int(self.off[-1])
        g = Graph()
        g.V, g.E = self.V, self.E
        g.off = self.off.ctypes.data_as(u64p)
        g.dst = self.dst.ctypes.data_as(u32p)
        g.w = self.w.ctypes.data_as(u32p) if self.w is not None else None
        g.stride = 1
        self.g = g


def _out(V, dt=np.uint32):
    return np.empty(V, dtype=dt)


def _call(name, *args):
    r = getattr(load(), name)(*args)
    if r < 0:
        raise RuntimeError(f"{name} failed ({r})")
    return r


def bfs_ref(csr, src, static_edges=None, part_edges=0, nthreads=0):
    out, st = _out(csr.V), Stats()
    S = csr.E if static_edges is None else static_edges
    _call("orc_bfs_ref", C.byref(csr.g), src, S, part_edges, nthreads, out.ctypes.data_as(u32p), C.byref(st))
    return out, st.as_dict()


def sssp_ref(csr, src, static_edges=None, part_edges=0, nthreads=0):
    out, st = _out(csr.V), Stats()
    S = csr.E if static_edges is None else static_edges
    _call("orc_sssp_ref", C.byref(csr.g), src, S, part_edges, nthreads, out.ctypes.data_as(u32p), C.byref(st))
    return out, st.as_dict()


def cc_ref(csr, static_edges=None, part_edges=0, nthreads=0):
    out, st = _out(csr.V), Stats()
    S = csr.E if static_edges is None else static_edges
    _call("orc_cc_ref", C.byref(csr.g), S, part_edges, nthreads, out.ctypes.data_as(u32p), C.byref(st))
    return out, st.as_dict()


def pr_ref(csr_in, outdeg, damping=0.85, threshold=0.001, max_iter=0):
    out, st = _out(csr_in.V, np.float64), Stats()
    od = np.ascontiguousarray(outdeg, dtype=np.uint32)
    _call("orc_pr_ref", C.byref(csr_in.g), od.ctypes.data_as(u32p), damping, threshold, max_iter,
          out.ctypes.data_as(f64p), C.byref(st))
    return out, st.as_dict()


def pr_norm(csr_in, outdeg, damping=0.85, iters=20):
    out = _out(csr_in.V, np.float64)
    res = np.empty(max(iters, 1), dtype=np.float64)
    od = np.ascontiguousarray(outdeg, dtype=np.uint32)
    _call("orc_pr_norm", C.byref(csr_in.g), od.ctypes.data_as(u32p), damping, iters, out.ctypes.data_as(f64p),
          res.ctypes.data_as(f64p))
    return out, res[:iters]


def bfs_queue(csr, src):
    out = _out(csr.V)
    _call("orc_bfs_queue", C.byref(csr.g), src, out.ctypes.data_as(u32p))
    return out


def dijkstra(csr, src):
    out = _out(csr.V)
    _call("orc_dijkstra", C.byref(csr.g), src, out.ctypes.data_as(u32p))
    return out


def cc_minreach(csr):
    out = _out(csr.V)
    _call("orc_cc_minreach", C.byref(csr.g), out.ctypes.data_as(u32p))
    return out


def cc_unionfind(csr):
    out = _out(csr.V)
    _call("orc_cc_unionfind", C.byref(csr.g), out.ctypes.data_as(u32p))
    return out


def plan_reference(off, edge_size, free_bytes, param_size, advise_rate=0.0, with_flags=False):
    off = np.ascontiguousarray(off, dtype=np.uint64)
    V = len(off) - 1
    p = Plan()
    flags = np.zeros(V, dtype=np.uint8) if with_flags else None
    _call("orc_plan_reference", V, int(off[-1]), off.ctypes.data_as(u64p), edge_size, free_bytes, param_size,
          advise_rate, C.byref(p), flags.ctypes.data_as(C.POINTER(C.c_uint8)) if with_flags else None)
    d = {n: getattr(p, n) for n, _ in p._fields_}
    return (d, flags) if with_flags else d


def split_reference(prefix, deg, part_size, cap=1 << 16):
    prefix = np.ascontiguousarray(prefix, dtype=np.uint64)
    deg = np.ascontiguousarray(deg, dtype=np.uint32)
    n = len(deg)
    total = int(prefix[n - 1] + deg[n - 1]) if n else 0
    s, c, e = np.zeros(cap, np.uint32), np.zeros(cap, np.uint32), np.zeros(cap, np.uint64)
    np_ = load().orc_split_reference(prefix.ctypes.data_as(u64p), deg.ctypes.data_as(u32p), n, total, part_size,
                                     s.ctypes.data_as(u32p), c.ctypes.data_as(u32p), e.ctypes.data_as(u64p), cap)
    if np_ < 0:
        return None
    return s[:np_], c[:np_], e[:np_]


def gather_ref(csr, vertices, nthreads=20, aos=None):
    """fillEdgeArrByMultiThread restated: copies the adjacency records of
    `vertices` into one buffer with `nthreads` threads.  aos: an (E, 2) u32
    array of {dst, w} records to copy instead of csr.dst (8-byte records).
    Returns (buffer, copy milliseconds)."""
    lst = np.ascontiguousarray(vertices, dtype=np.uint32)
    if aos is not None:
        aos = np.ascontiguousarray(aos, dtype=np.uint32)
        g = Graph()
        g.V, g.E, g.off = csr.V, csr.E, csr.g.off
        g.dst = aos.ctypes.data_as(u32p)
        g.w = None
        g.stride = 2
        words = 2
    else:
        g, words = csr.g, 1
    deg = (csr.off[lst.astype(np.int64) + 1] - csr.off[lst.astype(np.int64)]).sum() if len(lst) else 0
    out = np.empty(max(int(deg) * words, 1), dtype=np.uint32)
    ms = C.c_double()
    n = load().orc_gather_ref(C.byref(g), lst.ctypes.data_as(u32p), len(lst), nthreads, out.ctypes.data_as(u32p),
                              C.byref(ms))
    if n < 0:
        raise RuntimeError(f"orc_gather_ref failed ({n})")
    return out[:int(deg) * words], ms.value


class iteration_limit:
    """Context manager bounding the restated traversals to n iterations
    (timing samples; values are partial)."""

    def __init__(self, n):
        self.n = n

    def __enter__(self):
        load().orc_set_iteration_limit(self.n)

    def __exit__(self, *exc):
        load().orc_set_iteration_limit(0)


class edge_limit:
    """Context manager bounding the restated traversals and pr_norm to about
    n relaxed or pulled edges (single-thread timing samples at full size;
    values are partial)."""

    def __init__(self, n):
        self.n = n

    def __enter__(self):
        load().orc_set_edge_limit(self.n)

    def __exit__(self, *exc):
        load().orc_set_edge_limit(0)


def rmat_pairs(seed, scale, M, k, wmin=1, wmax=64, a=0.57, b=0.19, c=0.19):
    """Pairs and weights of k samples spread over the M samples of the
    library's R-MAT generator (u == v for a self-loop sample)."""
    u = np.empty(k, dtype=np.uint32)
    v = np.empty(k, dtype=np.uint32)
    w = np.empty(k, dtype=np.uint32)
    r = load().orc_rmat_pairs(seed, scale, a, b, c, M, wmin, wmax, k, u.ctypes.data_as(u32p),
                              v.ctypes.data_as(u32p), w.ctypes.data_as(u32p))
    if r != 0:
        raise ValueError("orc_rmat_pairs: bad parameters")
    return u, v, w


def last_pulled():
    return int(load().orc_last_pulled())


def set_num_threads(n):
    load().orc_set_num_threads(n)


class fault:
    """Context manager planting one of the oracle's deliberate mistakes."""

    def __init__(self, code):
        self.code = code

    def __enter__(self):
        load().orc_set_fault(self.code)

    def __exit__(self, *exc):
        load().orc_set_fault(0)
