"""GPU parity of the resident fast paths through the C ABI.

Direction-optimizing BFS and union-find CC run when every edge is on the
device and the adjacency is symmetric.  Their results must be bit-exact with
the oracle's BFS levels and min-reaching labels, and with the plain frontier
path of the same library (bfs_direction / cc_method forced).  The symmetry
flag that selects them is checked on graphs that miss one reverse edge or
differ in one multiplicity.
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


@pytest.fixture(scope="module")
def plain_ctx():
    """Context with the fast paths switched off (top-down BFS, label propagation)."""
    ctx = A.Context(bfs_direction=_abi.BFS_TOP_DOWN, cc_method=_abi.CC_LABEL_PROP)
    yield ctx
    ctx.close()


def host_csr(h):
    off = np.array(h.offsets(), dtype=np.uint64)
    dst = np.array(h.dst(), dtype=np.uint32)
    return off, dst


def _sym_graphs():
    out = []
    for scale, ef, seed in [(12, 16, 3), (14, 16, 4), (15, 8, 5)]:
        h = A.rmat(scale, ef, seed=seed)
        out.append((f"rmat_s{scale}_ef{ef}", *host_csr(h)))
    h = A.erdos_renyi(1 << 14, 1 << 17, seed=6)
    out.append(("er_2^14", *host_csr(h)))
    out.append(("path", *G.path_graph(3000)[:2]))
    out.append(("star", *G.star_graph(5000)[:2]))
    out.append(("two_components", *G.two_components(700, 1300)[:2]))
    # many small components whose smallest ids are not 0, plus isolated vertices
    e = []
    for c in range(200):
        b = 37 + c * 23
        e += [(b + i, b + (i * 7 + 3) % 20) for i in range(20)]
    out.append(("many_components", *G.csr_from_edges(37 + 200 * 23 + 11, e, symmetric=True)[:2]))
    return out


SYM = _sym_graphs()


@pytest.mark.parametrize("name,off,dst", SYM, ids=[s[0] for s in SYM])
@pytest.mark.parametrize("off32", [False, True])
def test_direction_opt_bfs_exact(gpu_ctx, plain_ctx, name, off, dst, off32):
    say("do-bfs", name, "off32", off32)
    V = len(off) - 1
    csr = O.CSR(off, dst)
    o = off.astype(np.uint32) if off32 else off
    deg = np.diff(off)
    srcs = sorted({0, int(np.argmax(deg)), int(np.nonzero(deg)[0][-1]) if deg.any() else 0, V - 1})
    with gpu_ctx.from_arrays(o, dst) as g, plain_ctx.from_arrays(o, dst) as gp:
        assert g.info()["symmetric"] == 1
        bu_edges = 0
        for s in srcs:
            lv, st = g.bfs(s)
            assert st["path"] == "direction_opt"
            ref = O.bfs_queue(csr, s)
            assert np.array_equal(lv, ref), (name, s)
            lp, stp = gp.bfs(s)
            assert stp["path"] == "frontier"
            assert np.array_equal(lp, lv)
            assert st["reached"] == stp["reached"] and st["reached_degree_sum"] == stp["reached_degree_sum"]
            bu_edges += st["kernels"]["bottom_up"]["edges"]
        if name.startswith("rmat") or name.startswith("er"):
            # wide middle levels: the bottom-up step must have run
            assert bu_edges > 0, name


@pytest.mark.parametrize("name,off,dst", SYM, ids=[s[0] for s in SYM])
def test_persistent_bfs_matches_host_stepped(name, off, dst):
    """The direction-optimizing BFS run as one persistent cooperative kernel
    (asc_config.bfs_persistent) and stepped level by level from the host (the
    default) give the oracle's levels in the same number of levels; the
    persistent one is a single launch."""
    say("persistent-bfs", name)
    csr = O.CSR(off, dst)
    deg = np.diff(off)
    srcs = sorted({0, int(np.argmax(deg)), len(off) - 2})
    with A.Context(profile=True, bfs_persistent=True) as c1, A.Context(profile=True) as c2, \
            c1.from_arrays(off, dst) as g1, c2.from_arrays(off, dst) as g2:
        for s in srcs:
            lv1, st1 = g1.bfs(s)
            lv2, st2 = g2.bfs(s)
            ref = O.bfs_queue(csr, s)
            assert np.array_equal(lv1, ref) and np.array_equal(lv2, ref), (name, s)
            assert st1["kernels"]["bfs_persistent"]["launches"] == 1
            assert st2["kernels"]["bfs_persistent"]["launches"] == 0
            assert st1["iterations"] == st2["iterations"]
            assert st1["reached"] == st2["reached"] == int((ref != len(off)).sum())


@pytest.mark.parametrize("name,off,dst", SYM, ids=[s[0] for s in SYM])
@pytest.mark.parametrize("off32", [False, True])
def test_union_find_cc_exact(gpu_ctx, plain_ctx, name, off, dst, off32):
    say("uf-cc", name, "off32", off32)
    csr = O.CSR(off, dst)
    o = off.astype(np.uint32) if off32 else off
    ref = O.cc_minreach(csr)
    with gpu_ctx.from_arrays(o, dst) as g, plain_ctx.from_arrays(o, dst) as gp:
        for rep in range(2):   # repeated runs on the same state
            lab, st = g.cc()
            assert st["path"] == "union_find"
            assert np.array_equal(lab, ref), (name, rep)
        lp, stp = gp.cc()
        assert stp["path"] == "frontier"
        assert np.array_equal(lp, lab)


def test_weighted_symmetric_graph_uses_fast_paths(gpu_ctx):
    h = A.rmat(13, 16, seed=8, weighted=True)
    off, dst = host_csr(h)
    w = np.array(h.weights(), dtype=np.uint32)
    csr = O.CSR(off, dst, w)
    with gpu_ctx.from_arrays(off, G.wedges(dst, w)) as g:
        assert g.info()["symmetric"] == 1
        lv, st = g.bfs(0)
        assert st["path"] == "direction_opt"
        assert np.array_equal(lv, O.bfs_queue(csr, 0))
        lab, st = g.cc()
        assert st["path"] == "union_find"
        assert np.array_equal(lab, O.cc_minreach(csr))
        d, st = g.sssp(0)
        assert st["path"] == "frontier"
        assert np.array_equal(d, O.dijkstra(csr, 0))


def test_symmetry_flag_negative_cases(gpu_ctx):
    off, dst, _ = G.random_graph(3000, 12000, 11, symmetric=True)
    edges = [(u, int(v)) for u in range(len(off) - 1) for v in dst[off[u]:off[u + 1]]]
    # one reverse edge missing
    missing = [e for e in edges if e != edges[len(edges) // 2]]
    o1, d1, _ = G.csr_from_edges(3000, missing, symmetric=False)
    # one edge doubled without its reverse (multiplicity differs)
    doubled = edges + [edges[7]]
    o2, d2, _ = G.csr_from_edges(3000, doubled, symmetric=False, dedup=False)
    # a directed random graph
    o3, d3, _ = G.random_graph(3000, 12000, 12, symmetric=False)
    for name, o, d in [("missing", o1, d1), ("doubled", o2, d2), ("directed", o3, d3)]:
        csr = O.CSR(o, d)
        with gpu_ctx.from_arrays(o, d) as g:
            assert g.info()["symmetric"] == 0, name
            lv, st = g.bfs(0)
            assert st["path"] == "frontier"
            assert np.array_equal(lv, O.bfs_queue(csr, 0)), name
            lab, st = g.cc()
            assert st["path"] == "frontier"
            assert np.array_equal(lab, O.cc_minreach(csr)), name
    # the doubled edge with its reverse doubled too is symmetric again
    o4, d4, _ = G.csr_from_edges(3000, doubled + [(edges[7][1], edges[7][0])], symmetric=False, dedup=False)
    with gpu_ctx.from_arrays(o4, d4) as g:
        assert g.info()["symmetric"] == 1


def test_partially_resident_graph_keeps_frontier_path():
    off, dst, _ = G.random_graph(5000, 20000, 13, symmetric=True)
    csr = O.CSR(off, dst)
    E = int(off[-1])
    with A.Context(edge_budget_bytes=4 * (E // 3), part_bytes=4 * 4096) as ctx, ctx.from_arrays(off, dst) as g:
        assert g.info()["symmetric"] == -1
        lv, st = g.bfs(0)
        assert st["path"] == "frontier"
        assert np.array_equal(lv, O.bfs_queue(csr, 0))
        lab, st = g.cc()
        assert st["path"] == "frontier"
        assert np.array_equal(lab, O.cc_minreach(csr))


def test_zero_based_convention_on_fast_path():
    off, dst, _ = G.two_components(500, 300)
    csr = O.CSR(off, dst)
    with A.Context(value_convention=_abi.VALUES_ZERO_BASED) as ctx, ctx.from_arrays(off, dst) as g:
        lv, st = g.bfs(0)
        assert st["path"] == "direction_opt"
        ref = O.bfs_queue(csr, 0).astype(np.int64)
        V = len(off) - 1
        exp = np.where(ref == V + 1, 0xFFFFFFFF, ref - 1).astype(np.uint32)
        assert np.array_equal(lv, exp)
