"""MI355X graph traversal: BFS, SSSP, CC and PageRank over binary CSR.

Python host interface over the C ABI of libascetic.so (include/ascetic.h).
The module-level functions `bfs_opt`, `cc_opt`, `sssp_opt` and `pr_opt`
keep the names and argument meaning of the reference drivers
(CalculateOpt.cuh:11-14: path, source, adviseRate with <= 0 meaning the
automatic rule) and return the result arrays the reference only prints
summaries of.  `Context` / `Graph` expose the full interface.

Everything here is host plumbing: the traversal runs in the HIP kernels of
libascetic.so, and a missing library raises instead of falling back.
"""
import ctypes as C

import numpy as np

from . import _abi
from ._abi import AscError, check  # noqa: F401

__all__ = [
    "Context", "Graph", "HostGraph", "generate", "read_graph", "rmat", "erdos_renyi", "write_graph",
    "plan_partition", "split_parts", "device_count", "bfs_opt", "cc_opt", "sssp_opt", "pr_opt",
    "AscError",
]


def _u32p(a):
    return a.ctypes.data_as(_abi.u32p)


def _u64p(a):
    return a.ctypes.data_as(_abi.u64p)


def device_count():
    return _abi.load().asc_device_count()


class HostGraph:
    """A CSR held in host memory by the library (generator output)."""

    def __init__(self, h):
        self._h = h
        self._owned = True

    @property
    def num_vertices(self):
        return self._h.num_vertices

    @property
    def num_edges(self):
        return self._h.num_edges

    @property
    def kind(self):
        return self._h.kind

    def offsets(self):
        """Zero-copy numpy view of the V+1 offsets."""
        return np.ctypeslib.as_array(self._h.off, shape=(self._h.num_vertices + 1,))

    def edge_words(self):
        """Zero-copy view of the edge array (E u32, or E x 2 {dst, w})."""
        E = self._h.num_edges
        if self._h.kind == _abi.KIND_WCSR:
            return np.ctypeslib.as_array(self._h.edges, shape=(max(E, 1), 2))[:E]
        return np.ctypeslib.as_array(self._h.edges, shape=(max(E, 1),))[:E]

    def dst(self):
        w = self.edge_words()
        return w[:, 0] if w.ndim == 2 else w

    def weights(self):
        w = self.edge_words()
        return w[:, 1] if w.ndim == 2 else None

    def out_degree(self):
        if not self._h.out_degree:
            return None
        return np.ctypeslib.as_array(self._h.out_degree, shape=(self._h.num_vertices,))

    def view(self):
        v = _abi.CsrView()
        _abi.load().asc_host_csr_view(C.byref(self._h), C.byref(v))
        return v

    def write(self, path, fmt):
        check(_abi.load().asc_write_graph(path.encode(), fmt, C.byref(self.view())), "asc_write_graph")

    def free(self):
        if self._owned and self._h.off:
            _abi.load().asc_host_csr_free(C.byref(self._h))
        self._owned = False

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def generate(type=0, scale=10, num_vertices=0, num_samples=0, edge_factor=16, a=0.57, b=0.19, c=0.19,
             seed=1, symmetrize=True, weighted=False, wmin=1, wmax=64, csc=False, threads=0, device=None):
    """Synthetic graph in host memory.  device: a Context whose GPU builds
    it (asc_generate_device, the same CSR bit for bit); None builds it on
    the host threads."""
    lib = _abi.load()
    p = _abi.GenParams()
    lib.asc_gen_params_init(C.byref(p))
    p.type, p.scale, p.num_vertices, p.num_samples = type, scale, num_vertices, num_samples
    p.edge_factor, p.a, p.b, p.c, p.seed = edge_factor, a, b, c, seed
    p.symmetrize, p.weighted, p.wmin, p.wmax = int(symmetrize), int(weighted), wmin, wmax
    p.csc, p.threads = int(csc), threads
    h = _abi.HostCsr()
    if device is not None:
        check(lib.asc_generate_device(device._h, C.byref(p), C.byref(h)), "asc_generate_device")
    else:
        check(lib.asc_generate(C.byref(p), C.byref(h)), "asc_generate")
    return HostGraph(h)


def read_graph(path, fmt=_abi.FMT_AUTO):
    """Reads a graph file into host memory (no device needed)."""
    h = _abi.HostCsr()
    check(_abi.load().asc_read_graph(path.encode(), fmt, C.byref(h)), f"asc_read_graph({path})")
    return HostGraph(h)


def read_edge_list(path, num_vertices=0, symmetrize=False, dedup=True, keep_self_loops=False, one_based=False,
                   weighted=False, csc=False, wmin=1, wmax=64, seed=1, threads=0):
    """Text / Matrix Market / .toB edge list -> HostGraph (sorted adjacency)."""
    lib = _abi.load()
    o = _abi.EdgeListOpts()
    lib.asc_edge_list_opts_init(C.byref(o))
    o.num_vertices, o.symmetrize, o.dedup, o.keep_self_loops = num_vertices, int(symmetrize), int(dedup), \
        int(keep_self_loops)
    o.one_based, o.weighted, o.csc, o.wmin, o.wmax, o.seed, o.threads = int(one_based), int(weighted), int(csc), \
        wmin, wmax, seed, threads
    h = _abi.HostCsr()
    check(lib.asc_read_edge_list(str(path).encode(), C.byref(o), C.byref(h)), f"asc_read_edge_list({path})")
    return HostGraph(h)


def rmat(scale, edge_factor=16, seed=1, **kw):
    return generate(type=0, scale=scale, edge_factor=edge_factor, seed=seed, **kw)


def erdos_renyi(num_vertices, num_samples, seed=1, **kw):
    return generate(type=1, num_vertices=num_vertices, num_samples=num_samples, seed=seed, **kw)


def _view_from_arrays(off, edges, kind, out_degree=None):
    """CsrView over numpy arrays (u64 or u32 offsets); the caller sets
    num_vertices and has_sentinel."""
    v = _abi.CsrView()
    if off.dtype == np.uint64:
        v.off64 = _u64p(off)
    elif off.dtype == np.uint32:
        v.off32 = _u32p(off)
    else:
        raise TypeError("offsets must be uint32 or uint64")
    v.num_edges = edges.shape[0]
    v.kind = kind
    v.edges = _u32p(edges)
    if out_degree is not None:
        v.out_degree = _u32p(out_degree)
    return v


def write_graph(path, fmt, off, edges, kind, out_degree=None):
    """Writes arrays in one of the binary layouts; `off` holds V entries
    (no sentinel), as the files do."""
    off = np.ascontiguousarray(off)
    edges = np.ascontiguousarray(edges, dtype=np.uint32)
    v = _view_from_arrays(off, edges, kind, out_degree)
    v.num_vertices = len(off)
    check(_abi.load().asc_write_graph(path.encode(), fmt, C.byref(v)), "asc_write_graph")


def plan_partition(off, edge_size, free_bytes, param_size, advise_rate=0.0, edge_budget_bytes=0, part_bytes=0):
    off = np.ascontiguousarray(off, dtype=np.uint64)
    V = len(off) - 1
    p = _abi.Plan()
    check(_abi.load().asc_plan_partition(V, int(off[-1]), _u64p(off), edge_size, free_bytes, param_size,
                                         advise_rate, edge_budget_bytes, part_bytes, C.byref(p)),
          "asc_plan_partition")
    return {name: getattr(p, name) for name, _ in p._fields_}


def split_parts(prefix, part_edges):
    prefix = np.ascontiguousarray(prefix, dtype=np.uint64)
    n = len(prefix) - 1
    count = C.c_uint64()
    lib = _abi.load()
    check(lib.asc_split_parts(_u64p(prefix), n, part_edges, None, 0, C.byref(count)), "asc_split_parts")
    out = np.zeros(count.value + 1, dtype=np.uint64)
    check(lib.asc_split_parts(_u64p(prefix), n, part_edges, _u64p(out), len(out), C.byref(count)),
          "asc_split_parts")
    return out


class Context:
    """One HIP device plus the library's streams and gather threads.  The
    keyword arguments are the fields of asc_config (include/ascetic.h); e.g.
    stream_pack selects the streamed record form: 1 packed (4 bytes when id and
    weight share a word, else 5: a u32 id plus a weight byte), 2 always the
    5-byte form, 0 the 8-byte {dst, w} records."""

    def __init__(self, device=0, advise_rate=0.0, edge_budget_bytes=0, part_bytes=0, device_mem_limit=0,
                 gather_threads=0, stream_mode=_abi.STREAM_GATHER, value_convention=_abi.VALUES_REFERENCE,
                 heavy_threshold=0, pr_group=0, sssp_delta=0, verbose=False, profile=False, stream_pack=1,
                 bfs_direction=_abi.BFS_AUTO, cc_method=_abi.CC_AUTO, dense_threshold=0.0, stream_overlap=None,
                 bfs_persistent=False, release_host_edges=False, part_buffers=0, max_pinned_bytes=0):
        lib = _abi.load()
        cfg = _abi.Config()
        lib.asc_config_init(C.byref(cfg))
        cfg.device, cfg.advise_rate, cfg.edge_budget_bytes = device, advise_rate, edge_budget_bytes
        cfg.part_bytes, cfg.device_mem_limit, cfg.gather_threads = part_bytes, device_mem_limit, gather_threads
        cfg.stream_mode, cfg.value_convention = stream_mode, value_convention
        cfg.heavy_threshold, cfg.pr_group, cfg.sssp_delta, cfg.verbose = (
            heavy_threshold, pr_group, sssp_delta, int(verbose))
        cfg.profile = int(profile)
        cfg.stream_pack = int(stream_pack)
        cfg.bfs_direction, cfg.cc_method = bfs_direction, cc_method
        cfg.dense_threshold = dense_threshold
        if stream_overlap is not None:
            cfg.stream_overlap = int(stream_overlap)
        cfg.bfs_persistent = int(bool(bfs_persistent))
        cfg.release_host_edges = int(bool(release_host_edges))
        cfg.part_buffers = int(part_buffers)
        cfg.max_pinned_bytes = int(max_pinned_bytes)
        h = C.c_void_p()
        check(lib.asc_ctx_create(C.byref(cfg), C.byref(h)), "asc_ctx_create")
        self._h = h
        self.value_convention = value_convention

    def close(self):
        if self._h:
            _abi.load().asc_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def synchronize(self):
        check(_abi.load().asc_ctx_synchronize(self._h), "asc_ctx_synchronize")

    def copy_probe(self, nbytes):
        """One plain device copy of nbytes (counter calibration); returns ms."""
        ms = C.c_double()
        check(_abi.load().asc_copy_probe(self._h, nbytes, C.byref(ms)), "asc_copy_probe")
        return ms.value

    def measure_link(self, nbytes=256 << 20, reps=8):
        """Pinned host <-> device copy bandwidth in GB/s: (h2d, d2h)."""
        h2d, d2h = C.c_double(), C.c_double()
        check(_abi.load().asc_measure_link(self._h, nbytes, reps, C.byref(h2d), C.byref(d2h)), "asc_measure_link")
        return h2d.value, d2h.value

    def load(self, path, fmt=_abi.FMT_AUTO):
        h = C.c_void_p()
        check(_abi.load().asc_graph_load(self._h, path.encode(), fmt, C.byref(h)), f"asc_graph_load({path})")
        return Graph(self, h)

    def adopt(self, host_graph, offset_bytes=8):
        """Moves a generated HostGraph into a device graph (no copy)."""
        h = C.c_void_p()
        check(_abi.load().asc_graph_adopt(self._h, C.byref(host_graph._h), offset_bytes, C.byref(h)),
              "asc_graph_adopt")
        host_graph._owned = False
        return Graph(self, h)

    def borrow_arrays(self, off, edges, kind=None, out_degree=None):
        """Graph over the caller's numpy arrays without copying them
        (asc_graph_borrow_csr); the Graph keeps references, and the arrays
        must not change while it is open.  off: V+1 offsets (u64 borrowed,
        u32 widened into a copy); edges: u32 [E] or u32 [E, 2] ({dst, w})."""
        off = np.ascontiguousarray(off)
        edges = np.ascontiguousarray(edges, dtype=np.uint32)
        if kind is None:
            kind = _abi.KIND_WCSR if edges.ndim == 2 else (_abi.KIND_CSC if out_degree is not None
                                                            else _abi.KIND_CSR)
        if out_degree is not None:
            out_degree = np.ascontiguousarray(out_degree, dtype=np.uint32)
        v = _view_from_arrays(off, edges, kind, out_degree)
        v.num_vertices = len(off) - 1
        v.has_sentinel = 1
        h = C.c_void_p()
        check(_abi.load().asc_graph_borrow_csr(self._h, C.byref(v), C.byref(h)), "asc_graph_borrow_csr")
        g = Graph(self, h)
        g._borrowed = (off, edges, out_degree)
        return g

    def generate_graph(self, type=0, scale=10, num_vertices=0, num_samples=0, edge_factor=16, a=0.57, b=0.19,
                       c=0.19, seed=1, symmetrize=True, wmin=1, wmax=64, offset_bytes=8):
        """Weighted synthetic graph generated straight into a device graph
        (asc_graph_generate): static prefix on the GPU, packed on-demand
        edges in pinned host memory, no host {dst, w} records."""
        lib = _abi.load()
        p = _abi.GenParams()
        lib.asc_gen_params_init(C.byref(p))
        p.type, p.scale, p.num_vertices, p.num_samples = type, scale, num_vertices, num_samples
        p.edge_factor, p.a, p.b, p.c, p.seed = edge_factor, a, b, c, seed
        p.symmetrize, p.weighted, p.wmin, p.wmax = int(symmetrize), 1, wmin, wmax
        p.offset_bytes = offset_bytes
        h = C.c_void_p()
        check(lib.asc_graph_generate(self._h, C.byref(p), C.byref(h)), "asc_graph_generate")
        return Graph(self, h)

    def from_arrays(self, off, edges, kind=None, out_degree=None):
        """Graph from numpy arrays: off has V+1 entries (u32 or u64), edges is
        u32 [E] or u32 [E, 2] ({dst, w})."""
        off = np.ascontiguousarray(off)
        edges = np.ascontiguousarray(edges, dtype=np.uint32)
        if kind is None:
            kind = _abi.KIND_WCSR if edges.ndim == 2 else (_abi.KIND_CSC if out_degree is not None
                                                            else _abi.KIND_CSR)
        if out_degree is not None:
            out_degree = np.ascontiguousarray(out_degree, dtype=np.uint32)
        v = _view_from_arrays(off, edges, kind, out_degree)
        v.num_vertices = len(off) - 1
        v.has_sentinel = 1
        h = C.c_void_p()
        check(_abi.load().asc_graph_from_csr(self._h, C.byref(v), C.byref(h)), "asc_graph_from_csr")
        return Graph(self, h)


class Graph:
    def __init__(self, ctx, h):
        self.ctx = ctx
        self._h = h

    def close(self):
        """Frees the graph; refused (AscError, ERR_STATE) while a server over
        it is open, and the handle then stays valid."""
        if self._h:
            check(_abi.load().asc_graph_destroy(self._h), "asc_graph_destroy")
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def info(self):
        i = _abi.GraphInfo()
        check(_abi.load().asc_graph_get_info(self._h, C.byref(i)), "asc_graph_get_info")
        return {name: getattr(i, name) for name, _ in i._fields_}

    @property
    def num_vertices(self):
        return self.info()["num_vertices"]

    def replan(self, edge_budget_bytes=0, advise_rate=0.0):
        check(_abi.load().asc_graph_replan(self._h, edge_budget_bytes, advise_rate), "asc_graph_replan")

    def autotune(self, budget_ms=30000.0):
        """Sweeps the tunables on this graph; returns the chosen values."""
        t = _abi.Tuning()
        check(_abi.load().asc_autotune(self._h, budget_ms, C.byref(t)), "asc_autotune")
        self._tuning = t
        return {n: getattr(t, n) for n, _ in t._fields_ if n != "pad"}

    def save_tuning(self, path):
        t = getattr(self, "_tuning", None)
        if t is None:
            raise RuntimeError("autotune() first")
        check(_abi.load().asc_tuning_save(path.encode(), self._h, C.byref(t)), "asc_tuning_save")

    def load_tuning(self, path):
        t = _abi.Tuning()
        check(_abi.load().asc_tuning_load(path.encode(), self._h, C.byref(t)), "asc_tuning_load")
        self._tuning = t
        return {n: getattr(t, n) for n, _ in t._fields_ if n != "pad"}

    def apply_tuning(self, **values):
        t = _abi.Tuning()
        for k, v in values.items():
            setattr(t, k, v)
        check(_abi.load().asc_tuning_apply(self._h, C.byref(t)), "asc_tuning_apply")

    def _run(self, fn, *args, output=True):
        out = np.empty(self.num_vertices, dtype=np.uint32) if output else None
        st = _abi.Stats()
        check(fn(self._h, *args, _u32p(out) if output else None, C.byref(st)), fn.__name__)
        return out, st.as_dict()

    def bfs(self, source=0, output=True):
        return self._run(_abi.load().asc_bfs, source, output=output)

    def sssp(self, source=0, output=True):
        return self._run(_abi.load().asc_sssp, source, output=output)

    def cc(self, output=True):
        return self._run(_abi.load().asc_cc, output=output)

    def pagerank(self, mode=_abi.PR_NORMALISED, damping=0.85, iterations=20, threshold=0.001,
                 max_iterations=0, method=_abi.PR_AUTO, tile_src_bits=0, tile_dst_bits=0):
        lib = _abi.load()
        o = _abi.PrOpts()
        lib.asc_pr_opts_init(C.byref(o))
        o.mode, o.damping, o.iterations, o.threshold, o.max_iterations = (
            mode, damping, iterations, threshold, max_iterations)
        o.method = method
        o.tile_src_bits, o.tile_dst_bits = tile_src_bits, tile_dst_bits
        V = self.num_vertices
        st = _abi.Stats()
        if mode == _abi.PR_NORMALISED:
            r = np.empty(V, dtype=np.float32)
            res = np.empty(max(iterations, 1), dtype=np.float64)
            check(lib.asc_pr(self._h, C.byref(o), r.ctypes.data_as(C.POINTER(C.c_float)), None,
                             res.ctypes.data_as(C.POINTER(C.c_double)), C.byref(st)), "asc_pr")
            return r, res, st.as_dict()
        r = np.empty(V, dtype=np.float64)
        check(lib.asc_pr(self._h, C.byref(o), None, r.ctypes.data_as(C.POINTER(C.c_double)), None,
                         C.byref(st)), "asc_pr")
        return r, None, st.as_dict()


class Server:
    """Continuous batching of BFS / SSSP queries over a resident graph.

    submit() is thread-safe and returns a ticket; when `out` is given (a
    uint32 array of V entries) it is filled before the query completes.
    wait(ticket) blocks until completion and returns the query's stats."""

    BFS = _abi.QUERY_BFS
    SSSP = _abi.QUERY_SSSP

    def __init__(self, graph, lanes=64, state_budget_bytes=0, queue_capacity=4096, bfs_direction=_abi.BFS_AUTO,
                 step_graphs=True):
        lib = _abi.load()
        o = _abi.ServerOpts()
        lib.asc_server_opts_init(C.byref(o))
        o.lanes, o.state_budget_bytes, o.queue_capacity = lanes, state_budget_bytes, queue_capacity
        o.bfs_direction = bfs_direction
        o.step_graphs = int(bool(step_graphs))
        h = C.c_void_p()
        check(lib.asc_server_create(graph._h, C.byref(o), C.byref(h)), "asc_server_create")
        self._h = h
        self.graph = graph
        self._V = graph.num_vertices
        self._outs = {}

    def submit(self, kind, source, out=None):
        t = C.c_uint64()
        ptr = None
        if out is not None:
            # the library writes V values into the buffer from its own thread
            if not (isinstance(out, np.ndarray) and out.dtype == np.uint32 and out.flags["C_CONTIGUOUS"]
                    and out.flags["WRITEABLE"] and out.size >= self._V):
                raise ValueError(f"out must be a writeable C-contiguous uint32 array of at least {self._V} entries")
            ptr = _u32p(out)
        check(_abi.load().asc_submit(self._h, kind, source, ptr, C.byref(t)), "asc_submit")
        if out is not None:
            self._outs[t.value] = out  # keep the buffer alive until completion
        return t.value

    def wait(self, ticket):
        qs = _abi.QStats()
        check(_abi.load().asc_wait(self._h, ticket, C.byref(qs)), "asc_wait")
        self._outs.pop(ticket, None)
        return {n: getattr(qs, n) for n, _ in qs._fields_}

    def drain(self):
        check(_abi.load().asc_server_drain(self._h), "asc_server_drain")

    def stats(self):
        s = _abi.ServerStats()
        check(_abi.load().asc_server_get_stats(self._h, C.byref(s)), "asc_server_get_stats")
        return {n: getattr(s, n) for n, _ in s._fields_}

    def close(self):
        if self._h:
            _abi.load().asc_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ---- reference-named drivers (CalculateOpt.cuh:11-14) ----------------------

def _driver(path, advise_rate, fmt, verbose, run):
    with Context(advise_rate=advise_rate if advise_rate > 0 else 0.0, verbose=verbose) as ctx:
        with ctx.load(path, fmt) as g:
            return run(g)


def bfs_opt(path, source_node=0, advise_rate=0.0, verbose=True):
    """BFS levels (source = 1, unreached = V + 1) of the graph file at path."""
    return _driver(path, advise_rate, _abi.FMT_AUTO, verbose, lambda g: g.bfs(source_node))


def sssp_opt(path, source_node=0, advise_rate=0.0, verbose=True):
    """SSSP distances (source = 1, unreached = V + 1) of a BWCSR file."""
    return _driver(path, advise_rate, _abi.FMT_AUTO, verbose, lambda g: g.sssp(source_node))


def cc_opt(path, advise_rate=0.0, verbose=True):
    """Min-label propagation labels of a BCSR file."""
    return _driver(path, advise_rate, _abi.FMT_AUTO, verbose, lambda g: g.cc())


def pr_opt(path, advise_rate=0.0, verbose=True, mode=_abi.PR_REFERENCE):
    """PageRank of a BCSC file; by default in the reference's FP64 active-set
    form (0.15 + 0.85 * sum, threshold 0.001)."""
    return _driver(path, advise_rate, _abi.FMT_BCSC, verbose, lambda g: g.pagerank(mode=mode))
