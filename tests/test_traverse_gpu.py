"""GPU parity: BFS / SSSP / CC through the C ABI against the oracle.

Integer results must be bit-exact.  Every case is checked with the graph
fully resident and with most of its edges on demand (small device budget,
small parts), in both streaming modes.
"""
import sys

import numpy as np
import pytest

import ascetic_amd as A
from ascetic_amd import _abi
from oracle import oracle as O
from tests import graphs as G

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_graph(ctx, off, dst, w=None, off32=False):
    off = np.asarray(off, np.uint32 if off32 else np.uint64)
    edges = dst if w is None else G.wedges(dst, w)
    return ctx.from_arrays(off, edges)


CASES = [
    ("path", lambda: G.path_graph(50)),
    ("star", lambda: G.star_graph(300)),
    ("two_components", lambda: G.two_components(40, 25)),
    ("random_sym", lambda: G.random_graph(5000, 20000, 1, symmetric=True)),
    ("random_dir", lambda: G.random_graph(5000, 20000, 2, symmetric=False)),
    ("hub", lambda: G.csr_from_edges(6000, [(0, i) for i in range(1, 6000)] +
                                     [(i, i + 1) for i in range(1, 5999)], symmetric=True)),
]


def weights_for(off, seed):
    return np.random.default_rng(seed).integers(1, 65, int(off[-1])).astype(np.uint32)


@pytest.mark.parametrize("name,mk", CASES)
@pytest.mark.parametrize("off32", [False, True])
def test_bfs_sssp_cc_resident(gpu_ctx, name, mk, off32):
    say("case", name, "off32", off32)
    off, dst, _ = mk()
    V = len(off) - 1
    # symmetric weights so a symmetric graph stays symmetric
    w = weights_for(off, 3)
    csr = O.CSR(off, dst, w)
    srcs = [0, int(np.argmax(np.diff(off))), V - 1]
    with make_graph(gpu_ctx, off, dst, off32=off32) as g:
        for s in srcs:
            lv, st = g.bfs(s)
            assert np.array_equal(lv, O.bfs_queue(csr, s)), (name, s)
            assert st["reached"] == int((lv != V + 1).sum())
        lab, _ = g.cc()
        assert np.array_equal(lab, O.cc_minreach(csr)), name
    with make_graph(gpu_ctx, off, dst, w, off32=off32) as g:
        for s in srcs:
            d, _ = g.sssp(s)
            assert np.array_equal(d, O.dijkstra(csr, s)), (name, s)
            lv, _ = g.bfs(s)
            assert np.array_equal(lv, O.bfs_queue(csr, s))


@pytest.mark.parametrize("mode", [_abi.STREAM_GATHER, _abi.STREAM_ZEROCOPY, _abi.STREAM_HYBRID])
@pytest.mark.parametrize("name,mk", CASES[3:])
def test_streaming_matches_resident(name, mk, mode):
    say("streaming", name, "mode", mode)
    off, dst, _ = mk()
    V = len(off) - 1
    w = weights_for(off, 5)
    csr = O.CSR(off, dst, w)
    E = int(off[-1])
    # static prefix ~ 1/3 of the edges, parts of 4096 edges (a hub spans parts)
    with A.Context(edge_budget_bytes=4 * (E // 3), part_bytes=4 * 4096, stream_mode=mode) as ctx:
        with make_graph(ctx, off, dst) as g:
            info = g.info()
            # (the hub case's first vertex already crosses the 4096-edge
            # floored prefix, so nothing but isolated vertices is static)
            assert info["static_vertices"] < V and info["ondemand_edges"] > 0
            for s in (0, int(np.argmax(np.diff(off)))):
                lv, st = g.bfs(s)
                assert np.array_equal(lv, O.bfs_queue(csr, s))
                assert st["streamed_edges"] > 0 or s == 0
            lab, st = g.cc()
            assert np.array_equal(lab, O.cc_minreach(csr))
            assert st["streamed_edges"] > 0
            # on-demand edges are counted under the device-buffer part kernel
            # or, when read from mapped host memory, under expand_part_host
            kp, kh = st["kernels"]["expand_part"], st["kernels"]["expand_part_host"]
            assert kp["edges"] + kh["edges"] == st["edges_traversed"] - st["static_edges_traversed"] > 0
            if mode == _abi.STREAM_ZEROCOPY:
                assert kp["edges"] == 0 and kh["edges"] > 0
    # stream_pack 1: 4-byte words; 2: u32 ids plus weight bytes; 0: 8-byte {dst, w} records
    for pack, rec in ((1, 4), (2, 5), (0, 8)):
        with A.Context(edge_budget_bytes=8 * (E // 3), part_bytes=8 * 4096, stream_mode=mode,
                       stream_pack=pack) as ctx:
            with make_graph(ctx, off, dst, w) as g:
                for s in (0, int(np.argmax(np.diff(off)))):
                    d, st = g.sssp(s)
                    assert np.array_equal(d, O.dijkstra(csr, s)), (pack, s)
                    if st["streamed_edges"]:
                        assert st["streamed_bytes"] == st["streamed_edges"] * rec
                    lv, _ = g.bfs(s)
                    assert np.array_equal(lv, O.bfs_queue(csr, s))


@pytest.mark.parametrize("thr", [1e-9, 0.5, 2.0])
@pytest.mark.parametrize("name,mk", CASES[3:])
def test_dense_chunk_copies(name, mk, thr):
    """On-demand chunks copied whole by DMA (threshold ~0: every active chunk;
    0.5: the default; 2: never) give the same values as the oracle, in the
    4-byte packed, split 5-byte, 8-byte and unweighted layouts."""
    say("dense", name, "threshold", thr)
    off, dst, _ = mk()
    w = weights_for(off, 6)
    csr = O.CSR(off, dst, w)
    E = int(off[-1])
    srcs = (0, int(np.argmax(np.diff(off))))
    dense = 0
    with A.Context(edge_budget_bytes=4 * (E // 3), part_bytes=4 * 4096, dense_threshold=thr) as ctx:
        with make_graph(ctx, off, dst) as g:
            for s in srcs:
                lv, st = g.bfs(s)
                assert np.array_equal(lv, O.bfs_queue(csr, s)), s
                assert st["dense_edges"] <= st["streamed_edges"]
                dense += st["dense_edges"]
            lab, st = g.cc()
            assert np.array_equal(lab, O.cc_minreach(csr))
            dense += st["dense_edges"]
    for pack, rec in ((1, 4), (2, 5), (0, 8)):
        with A.Context(edge_budget_bytes=8 * (E // 3), part_bytes=8 * 4096, stream_pack=pack,
                       dense_threshold=thr) as ctx:
            with make_graph(ctx, off, dst, w) as g:
                for s in srcs:
                    d, st = g.sssp(s)
                    assert np.array_equal(d, O.dijkstra(csr, s)), (pack, s)
                    assert st["streamed_bytes"] == st["streamed_edges"] * rec
                    dense += st["dense_edges"]
    if thr > 1:
        assert dense == 0
    elif thr < 1e-6:
        assert dense > 0


@pytest.mark.parametrize("every", [1, 3])
@pytest.mark.parametrize("pack", [1, 2])
@pytest.mark.parametrize("thr", [1e-9, 2.0])
@pytest.mark.parametrize("name,mk", CASES[3:])
def test_stream_overlap_exact(name, mk, thr, pack, every):
    """Resident steps between streamed chunks / parts (SSSP, CC; after every
    part or every third) give the oracle's values, with whole-chunk copies and
    with gathering only, in the 4-byte and the split 5-byte stream forms."""
    say("overlap", name, "threshold", thr, "pack", pack, "every", every)
    off, dst, _ = mk()
    w = weights_for(off, 8)
    csr = O.CSR(off, dst, w)
    E = int(off[-1])
    with A.Context(edge_budget_bytes=8 * (E // 3), part_bytes=8 * 2048, dense_threshold=thr,
                   stream_overlap=every, stream_pack=pack) as ctx:
        with make_graph(ctx, off, dst, w) as g:
            for s in (0, int(np.argmax(np.diff(off))), len(off) - 2):
                d, st = g.sssp(s)
                assert np.array_equal(d, O.dijkstra(csr, s)), (name, s)
            lab, _ = g.cc()
            assert np.array_equal(lab, O.cc_minreach(csr)), name
            lv, _ = g.bfs(0)
            assert np.array_equal(lv, O.bfs_queue(csr, 0)), name


@pytest.mark.parametrize("pack", [1, 2])
@pytest.mark.parametrize("mode", [_abi.STREAM_GATHER, _abi.STREAM_ZEROCOPY])
def test_everything_on_demand(gpu_ctx, mode, pack):
    """A device budget of one byte: every edge is on demand. In zero-copy mode
    kernels read the whole graph from host memory at the link, the form the
    reference's unified-memory ("Share") drivers take without page migration."""
    off, dst, _ = G.random_graph(3000, 12000, 9, symmetric=False)
    w = weights_for(off, 9)
    csr = O.CSR(off, dst, w)
    with A.Context(edge_budget_bytes=1, part_bytes=8 * 1024, stream_mode=mode, stream_pack=pack) as ctx:
        with make_graph(ctx, off, dst, w) as g:
            assert g.info()["static_vertices"] <= 1
            d, st = g.sssp(5)
            assert np.array_equal(d, O.dijkstra(csr, 5))
            assert st["parts"] >= 1
            lab, _ = g.cc()
            assert np.array_equal(lab, O.cc_minreach(csr))
            lv, _ = g.bfs(5)
            assert np.array_equal(lv, O.bfs_queue(csr, 5))


def _graph_with_heavy_chain(seed):
    """100000 vertices with 250000 random edges of weight 1-16, plus a chain
    of 40 vertices hung from vertex 0 by edges of weight 3000: the mean weight
    stays small while the chain's values pass 65535 (symmetric weights)."""
    rng = np.random.default_rng(seed)
    u = rng.integers(0, 100000, 250000)
    v = rng.integers(0, 100000, 250000)
    keep = u != v
    u, v = u[keep], v[keep]
    w = rng.integers(1, 17, len(u))
    chain = np.arange(100000, 100040)
    u = np.concatenate([u, [0], chain[:-1]])
    v = np.concatenate([v, chain[:1], chain[1:]])
    w = np.concatenate([w, np.full(40, 3000)])
    return G.csr_from_edges(100040, zip(u.tolist(), v.tolist()), w.tolist(), symmetric=True)


@pytest.mark.parametrize("kind,wmax", [("chain", 3000), ("random", 20000), ("random", 1 << 20)])
def test_resident_sssp_values_beyond_16_bits(gpu_ctx, kind, wmax):
    """Resident SSSP probes a 16-bit copy of the values before its atomics
    (0xFFFF: nothing known, the atomic is tried) where the values should fit
    16 bits.  With 100000 vertices a value can exceed 65535 (values are
    capped at V + 1): the heavy chain keeps the filter on with values on both
    sides of 65535; random weights up to 20000 or 2^20 turn it off."""
    say("resident sssp", kind, "weights up to", wmax)
    if kind == "chain":
        off, dst, w = _graph_with_heavy_chain(11)
    else:
        off, dst, _ = G.random_graph(100000, 250000, 11, symmetric=True)
        w = np.random.default_rng(12).integers(1, wmax + 1, int(off[-1])).astype(np.uint32)
    V = len(off) - 1
    csr = O.CSR(off, dst, w)
    with make_graph(gpu_ctx, off, dst, w) as g:
        assert g.info()["static_vertices"] == V
        for s in (0, int(np.argmax(np.diff(off)))):
            ref = O.dijkstra(csr, s)
            if wmax <= 20000:
                reached = ref[ref != V + 1]
                assert (reached < 0xFFFF).any() and (reached > 0xFFFF).any()
            d, _ = g.sssp(s)
            assert np.array_equal(d, ref), s


def test_value_conventions():
    off, dst, _ = G.two_components(5, 3)
    with A.Context(value_convention=_abi.VALUES_ZERO_BASED) as ctx:
        with make_graph(ctx, off, dst) as g:
            lv, _ = g.bfs(1)
    assert lv.tolist() == [1, 0, 1, 2, 3] + [0xFFFFFFFF] * 3


def test_errors(gpu_ctx):
    off, dst, _ = G.path_graph(10)
    with make_graph(gpu_ctx, off, dst) as g:
        with pytest.raises(A.AscError) as e:
            g.bfs(10)
        assert e.value.code == _abi.ERR_INVALID
        with pytest.raises(A.AscError) as e:
            g.sssp(0)
        assert e.value.code == _abi.ERR_UNSUPPORTED


def test_rmat_scale20_bfs_from_0(gpu_ctx):
    """R-MAT scale 20, edge factor 16, symmetrized, u32 offsets; BFS from 0."""
    say("rmat s20 bfs")
    h = A.rmat(20, 16, seed=1)
    off, dst = h.offsets().copy(), h.dst().copy()
    ref = O.bfs_queue(O.CSR(off, dst), 0)
    with gpu_ctx.adopt(h, offset_bytes=4) as g:
        for _ in range(2):
            lv, st = g.bfs(0)
            assert np.array_equal(lv, ref)
        assert st["reached"] == int((ref != len(off)).sum())


def test_rmat_scale18_sssp_budget40(gpu_ctx):
    """Weighted R-MAT (w in [1,64]), source = max-degree vertex, device edge
    budget 40% of the edge+weight bytes: the rest streams each iteration."""
    say("rmat s18 sssp budget 40%")
    h = A.rmat(18, 16, seed=2, weighted=True, wmin=1, wmax=64)
    off = h.offsets().copy()
    ew = h.edge_words().copy()
    E = int(off[-1])
    src = int(np.argmax(np.diff(off)))
    ref = O.dijkstra(O.CSR(off, ew[:, 0], ew[:, 1]), src)
    for mode in (_abi.STREAM_GATHER, _abi.STREAM_ZEROCOPY):
        with A.Context(edge_budget_bytes=int(0.4 * 8 * E), part_bytes=8 << 20, stream_mode=mode) as ctx:
            with ctx.from_arrays(off, ew) as g:
                d, st = g.sssp(src)
                assert np.array_equal(d, ref)
                assert st["streamed_edges"] > 0
