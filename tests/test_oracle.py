"""Pins the oracle (oracle/oracle.c) before it is trusted as the checker.

The reference repository ships no tests, fixtures or golden outputs for its
hot path, so the restatement of CalculateOpt.cu's iteration is pinned by:
textbook algorithms (queue BFS, Dijkstra, union-find), third-party
libraries (scipy.sparse.csgraph, networkx), closed forms on hand graphs,
brute-force path enumeration on tiny graphs, fixed points of PageRank, and
planted faults that each of these checks must catch.
"""
import itertools
import os

import networkx as nx
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse import csgraph

from oracle import oracle as O
from tests import graphs as G

pytestmark = []


def sp_matrix(off, dst, w, V):
    data = np.ones(len(dst)) if w is None else w.astype(np.float64)
    return sp.csr_matrix((data, dst.astype(np.int64), off.astype(np.int64)), shape=(V, V))


def scipy_levels(off, dst, V, src):
    d = csgraph.shortest_path(sp_matrix(off, dst, None, V), method="D", directed=True, indices=src,
                              unweighted=True)
    return np.where(np.isinf(d), V + 1, d + 1).astype(np.uint32)


def scipy_sssp(off, dst, w, V, src):
    d = csgraph.shortest_path(sp_matrix(off, dst, w, V), method="D", directed=True, indices=src)
    out = np.where(np.isinf(d), V + 1, d + 1)
    return np.minimum(out, V + 1).astype(np.uint32)


def scipy_cc_min(off, dst, V):
    n, lab = csgraph.connected_components(sp_matrix(off, dst, None, V), directed=False)
    mins = np.full(n, V, dtype=np.int64)
    np.minimum.at(mins, lab, np.arange(V))
    return mins[lab].astype(np.uint32)


# ---- closed forms ------------------------------------------------------------

def test_bfs_closed_forms():
    off, dst, _ = G.path_graph(7)
    lv, st = O.bfs_ref(O.CSR(off, dst), 0)
    assert lv.tolist() == [1, 2, 3, 4, 5, 6, 7]
    assert st["iterations"] == 7  # one level per vertex; the last level finds nothing
    lv, _ = O.bfs_ref(O.CSR(off, dst), 3)
    assert lv.tolist() == [4, 3, 2, 1, 2, 3, 4]
    off, dst, _ = G.star_graph(6)
    assert O.bfs_ref(O.CSR(off, dst), 0)[0].tolist() == [1, 2, 2, 2, 2, 2]
    assert O.bfs_ref(O.CSR(off, dst), 4)[0].tolist() == [2, 3, 3, 3, 1, 3]
    off, dst, _ = G.two_components(3, 2)
    assert O.bfs_ref(O.CSR(off, dst), 0)[0].tolist() == [1, 2, 3, 6, 6]  # unreached = V + 1


def test_sssp_closed_forms():
    # path with weights 5, 1, 7 (16 extra isolated vertices, V = 20):
    # distances 0, 5, 6, 13 -> 1, 6, 7, 14
    off, dst, w = G.csr_from_edges(20, [(0, 1), (1, 2), (2, 3)], [5, 1, 7], symmetric=True)
    assert O.sssp_ref(O.CSR(off, dst, w), 0)[0].tolist()[:4] == [1, 6, 7, 14]
    # triangle where the two-hop route is shorter than the direct edge
    off, dst, w = G.csr_from_edges(10, [(0, 1), (1, 2), (0, 2)], [1, 1, 5], symmetric=True)
    assert O.sssp_ref(O.CSR(off, dst, w), 0)[0].tolist()[:3] == [1, 2, 3]
    # the reference starts every vertex at V + 1 and relaxes only on `<`
    # (GraphMeta.cuh:440-446, gpu_kernels.cuh:499): a vertex whose distance
    # + 1 reaches V + 1 keeps the unreached value
    off, dst, w = G.csr_from_edges(4, [(0, 1), (1, 2), (2, 3)], [5, 1, 7], symmetric=True)
    assert O.sssp_ref(O.CSR(off, dst, w), 0)[0].tolist() == [1, 5, 5, 5]
    off, dst, w = G.csr_from_edges(4, [(0, 1), (1, 2)], [2, 1], symmetric=False)
    assert O.sssp_ref(O.CSR(off, dst, w), 0)[0].tolist() == [1, 3, 4, 5]
    assert O.dijkstra(O.CSR(off, dst, w), 0).tolist() == [1, 3, 4, 5]


def test_cc_closed_forms():
    off, dst, _ = G.two_components(3, 4)
    assert O.cc_ref(O.CSR(off, dst))[0].tolist() == [0, 0, 0, 3, 3, 3, 3]
    # directed: label = min id that reaches v (3 -> 1 -> 2, 0 isolated)
    off, dst, _ = G.csr_from_edges(4, [(3, 1), (1, 2)])
    assert O.cc_ref(O.CSR(off, dst))[0].tolist() == [0, 1, 1, 3]


# ---- brute force on tiny graphs ---------------------------------------------

def brute_paths(V, edges, weights, src):
    """Shortest simple-path length from src by exhaustive DFS."""
    adj = {u: [] for u in range(V)}
    for (u, v), w in zip(edges, weights):
        adj[u].append((v, w))
    best = [None] * V

    def dfs(u, d, seen):
        if best[u] is None or d < best[u]:
            best[u] = d
        for v, w in adj[u]:
            if v not in seen:
                dfs(v, d + w, seen | {v})

    dfs(src, 0, {src})
    return best


@pytest.mark.parametrize("seed", range(25))
def test_brute_force_tiny(seed):
    rng = np.random.default_rng(seed)
    V = int(rng.integers(2, 8))
    M = int(rng.integers(0, 14))
    pairs = [(int(a), int(b)) for a, b in zip(rng.integers(0, V, M), rng.integers(0, V, M)) if a != b]
    pairs = sorted(set(pairs))
    ws = [int(x) for x in rng.integers(1, 9, len(pairs))]
    off, dst, w = G.csr_from_edges(V, pairs, ws)
    csr = O.CSR(off, dst, w)
    for src in range(V):
        hops = brute_paths(V, pairs, [1] * len(pairs), src)
        dist = brute_paths(V, pairs, ws, src)
        exp_b = [V + 1 if h is None else h + 1 for h in hops]
        exp_s = [V + 1 if d is None else min(d + 1, V + 1) for d in dist]
        for S in (0, int(off[-1]) // 2, int(off[-1])):
            assert O.bfs_ref(csr, src, S, 2)[0].tolist() == exp_b
            assert O.sssp_ref(csr, src, S, 3)[0].tolist() == exp_s
    reach = np.eye(V, dtype=bool)
    for u, v in pairs:
        reach[u, v] = True
    for _ in range(V):
        reach = reach | (reach.astype(np.int64) @ reach.astype(np.int64) > 0)
    exp_cc = [int(np.nonzero(reach[:, v])[0].min()) for v in range(V)]
    assert O.cc_ref(csr)[0].tolist() == exp_cc
    assert O.cc_minreach(csr).tolist() == exp_cc


# ---- textbook and third-party cross checks ----------------------------------

@pytest.mark.parametrize("symmetric", [True, False])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_traversal_vs_libraries(symmetric, seed):
    V = 3000
    off, dst, w = G.random_graph(V, 9000, seed, symmetric=symmetric, weighted=True, wmax=64)
    csr = O.CSR(off, dst, w)
    E = int(off[-1])
    src = int(np.argmax(np.diff(off)))
    lv_sp = scipy_levels(off, dst, V, src)
    d_sp = scipy_sssp(off, dst, w, V, src)
    assert np.array_equal(O.bfs_queue(csr, src), lv_sp)
    assert np.array_equal(O.dijkstra(csr, src), d_sp)
    # the restatement, resident and with most edges on demand in small parts
    for S, P, nt in ((E, 0, 1), (E // 3, 257, 4), (0, 64, 8)):
        assert np.array_equal(O.bfs_ref(csr, src, S, P, nt)[0], lv_sp)
        assert np.array_equal(O.sssp_ref(csr, src, S, P, nt)[0], d_sp)
        lab = O.cc_ref(csr, S, P, nt)[0]
        assert np.array_equal(lab, O.cc_minreach(csr))
        if symmetric:
            assert np.array_equal(lab, scipy_cc_min(off, dst, V))
            assert np.array_equal(lab, O.cc_unionfind(csr))


def test_bfs_ref_nodesum_counts_reached():
    # BFS stays level-synchronous, so nodeSum = number of reached vertices
    off, dst, _ = G.random_graph(2000, 6000, 7)
    csr = O.CSR(off, dst)
    lv, st = O.bfs_ref(csr, 0, int(off[-1]) // 2, 100, 4)
    assert st["node_sum"] == int((lv != 2001).sum())
    assert st["iterations"] == int(lv[lv != 2001].max())


# ---- PageRank ------------------------------------------------------------------

def test_pr_ref_two_cycle_closed_form():
    # 0 <-> 1, out-degree 1: r_k = 1 - 0.5 * 0.85^k.  A vertex stops when
    # |r_{k+1} - r_k| <= 0.001 and keeps r_k (the un-stored t is dropped).
    off, src, _ = G.csr_from_edges(2, [(0, 1), (1, 0)])
    val, st = O.pr_ref(O.CSR(off, src), np.array([1, 1], np.uint32))
    rk = [0.5]
    while True:
        t = 0.15 + 0.85 * rk[-1]
        if abs(t - rk[-1]) <= 0.001:
            break
        rk.append(t)
    assert st["iterations"] == len(rk)
    assert val.tolist() == [rk[-1], rk[-1]]


def test_pr_ref_fixed_point():
    # with a tiny threshold every vertex converges to r = 0.15 + 0.85 P^T r
    V = 200
    off, dst, _ = G.random_graph(V, 900, 3, symmetric=False)
    deg = np.diff(off).astype(np.uint32)
    ioff, isrc = G.transpose(off, dst, V)
    od = np.maximum(deg, 1)
    P = sp_matrix(off, dst, None, V).multiply(1.0 / od[:, None]).tocsr()
    x = np.linalg.solve(np.eye(V) - 0.85 * P.T.toarray(), np.full(V, 0.15))
    val, _ = O.pr_ref(O.CSR(ioff, isrc), deg, threshold=1e-13, max_iter=5000)
    assert np.max(np.abs(val - x)) < 1e-9


@pytest.mark.parametrize("symmetric", [True, False])
def test_pr_norm_vs_networkx(symmetric):
    V = 500
    off, dst, _ = G.random_graph(V, 2000, 11, symmetric=symmetric)
    deg = np.diff(off).astype(np.uint32)
    g = nx.DiGraph()
    g.add_nodes_from(range(V))
    src = np.repeat(np.arange(V), deg)
    g.add_edges_from(zip(src.tolist(), dst.tolist()))
    ref = nx.pagerank(g, alpha=0.85, tol=1e-15, max_iter=10000)
    ref = np.array([ref[i] for i in range(V)])
    ioff, isrc = G.transpose(off, dst, V)
    r, res = O.pr_norm(O.CSR(ioff, isrc), deg, 0.85, 300)
    assert np.abs(r - ref).sum() < 1e-10
    # 20 iterations: the L1 error contracts at least by the damping factor
    r20, res20 = O.pr_norm(O.CSR(ioff, isrc), deg, 0.85, 20)
    err0 = np.abs(np.full(V, 1.0 / V) - ref).sum()
    assert np.abs(r20 - ref).sum() <= err0 * 0.85 ** 20 * 1.0001 + 1e-15
    assert abs(r20.sum() - 1.0) < 1e-12
    assert np.all(np.diff(res20) <= 1e-15)


# ---- planning ----------------------------------------------------------------

def test_plan_reference_hand_values():
    V = 1000
    off = (np.arange(V + 1) * 10).astype(np.uint64)  # every degree 10, E = 10000
    base = 13 * 4 * V
    # total = 10000 edges -> adviseK = (10 - 1)/9 = 1 -> max = 10000, floored to 8192
    p, flags = O.plan_reference(off, 4, base + 4 * 10000, 13, with_flags=True)
    assert p["total_gpu_size"] == 10000 and p["advise_k"] == 1.0
    assert p["max_partition_size"] == 8192
    assert p["static_vertices"] == 819 and flags[:819].all() and not flags[819:].any()
    assert p["overload_size"] == 10000 - 8190 and p["part_overload_size"] == 10000 - 8192
    # total = 5000 -> (10 - 2)/9 -> 0.888.. * 5000 = 4444 -> 4096
    p = O.plan_reference(off, 4, base + 4 * 5000, 13)
    assert p["max_partition_size"] == 4096
    # tiny memory -> negative rule -> 0.5
    p = O.plan_reference(off, 4, base + 4 * 500, 13)
    assert p["advise_k"] == 0.5 and p["max_partition_size"] == 0
    # explicit adviseK wins
    p = O.plan_reference(off, 4, base + 4 * 100000, 13, advise_rate=0.25)
    assert p["advise_k"] == 0.25 and p["max_partition_size"] == 8192  # 25000 -> E = 10000 -> 8192
    p = O.plan_reference(off, 4, base + 4 * 30000, 13, advise_rate=0.25)
    assert p["max_partition_size"] == 4096  # 7500 -> 4096


def test_split_reference_invariants_and_hang():
    rng = np.random.default_rng(5)
    deg = rng.integers(1, 50, 300).astype(np.uint32)
    prefix = np.concatenate([[0], np.cumsum(deg)[:-1]]).astype(np.uint64)
    total = int(deg.sum())
    s, c, e = O.split_reference(prefix, deg, 200)
    assert s[0] == 0 and int(c.sum()) == len(deg) and int(e.sum()) == total
    assert np.all(e <= 200)
    assert np.array_equal(s[1:], (s + c)[:-1])
    # a vertex larger than a part: the reference's loop makes no progress
    deg2 = deg.copy()
    deg2[10] = 500
    prefix2 = np.concatenate([[0], np.cumsum(deg2)[:-1]]).astype(np.uint64)
    assert O.split_reference(prefix2, deg2, 200) is None


# ---- planted faults: each must be caught by the checks above -----------------

def _fault_caught(code, check):
    with O.fault(code):
        try:
            check()
        except AssertionError:
            return True
    return False


def test_planted_faults_are_caught():
    V = 400
    off, dst, w = G.random_graph(V, 1500, 21, symmetric=True, weighted=True)
    csr = O.CSR(off, dst, w)
    E = int(off[-1])
    src = int(np.argmax(np.diff(off)))

    def bfs_check():
        assert np.array_equal(O.bfs_ref(csr, src)[0], scipy_levels(off, dst, V, src))

    def sssp_check():
        assert np.array_equal(O.sssp_ref(csr, src, E // 4, 50, 2)[0], scipy_sssp(off, dst, w, V, src))

    def cc_check():
        assert np.array_equal(O.cc_ref(csr)[0], scipy_cc_min(off, dst, V))

    offd, dstd, _ = G.random_graph(V, 1500, 22, symmetric=False)
    degd = np.diff(offd).astype(np.uint32)
    ioff, isrc = G.transpose(offd, dstd, V)

    def pr_ref_check():
        od = np.maximum(degd, 1)
        P = sp_matrix(offd, dstd, None, V).multiply(1.0 / od[:, None]).tocsr()
        x = np.linalg.solve(np.eye(V) - 0.85 * P.T.toarray(), np.full(V, 0.15))
        assert np.max(np.abs(O.pr_ref(O.CSR(ioff, isrc), degd, threshold=1e-13, max_iter=5000)[0] - x)) < 1e-9

    def pr_norm_check():
        g = nx.DiGraph()
        g.add_nodes_from(range(V))
        g.add_edges_from(zip(np.repeat(np.arange(V), degd).tolist(), dstd.tolist()))
        ref = nx.pagerank(g, alpha=0.85, tol=1e-15, max_iter=10000)
        ref = np.array([ref[i] for i in range(V)])
        assert np.abs(O.pr_norm(O.CSR(ioff, isrc), degd, 0.85, 300)[0] - ref).sum() < 1e-10

    def part_check():
        assert np.array_equal(O.bfs_ref(csr, src, 0, 37, 3)[0], scipy_levels(off, dst, V, src))

    # the checks pass without faults
    for chk in (bfs_check, sssp_check, cc_check, pr_ref_check, pr_norm_check, part_check):
        chk()
    assert _fault_caught(1, bfs_check)
    assert _fault_caught(2, sssp_check)
    assert _fault_caught(3, cc_check)
    assert _fault_caught(4, pr_ref_check)
    assert _fault_caught(5, pr_norm_check)
    assert _fault_caught(6, part_check)


def test_gather_ref_copies_adjacencies_in_list_order():
    """fillEdgeArrByMultiThread restated: the adjacency records of the listed
    vertices, concatenated in list order, whatever the thread count."""
    off, dst, w = G.random_graph(500, 3000, 41, symmetric=True, weighted=True)
    csr = O.CSR(off, dst, w)
    lst = np.array([7, 3, 499, 0, 250, 3], np.uint32)
    expect = np.concatenate([dst[off[v]:off[v + 1]] for v in lst])
    for nt in (1, 3, 20):
        got, ms = O.gather_ref(csr, lst, nt)
        assert np.array_equal(got, expect) and ms >= 0
    aos = G.wedges(dst, w)
    big = np.arange(0, 500, 2, dtype=np.uint32)          # >= 50 vertices: 20 threads
    got, _ = O.gather_ref(csr, big, 20, aos=aos)
    assert np.array_equal(got.reshape(-1, 2), np.concatenate([aos[off[v]:off[v + 1]] for v in big]))


def test_iteration_limit_bounds_the_restatement():
    """With a limit of k iterations, BFS on a path reaches exactly the first
    k + 1 vertices; without it, the whole path."""
    off, dst, _ = G.path_graph(20)
    csr = O.CSR(off, dst)
    with O.iteration_limit(3):
        lv, st = O.bfs_ref(csr, 0)
    assert st["iterations"] == 3
    assert lv[:4].tolist() == [1, 2, 3, 4] and (lv[4:] == 21).all()
    lv, st = O.bfs_ref(csr, 0)
    assert lv.tolist() == list(range(1, 21))


def test_edge_limit_bounds_the_restatement():
    """On one thread an edge limit of n stops the restated traversal and the
    FP64 power iteration after about n edges (the last vertex started is
    finished), and a limit of 0 leaves the whole run."""
    off, dst = G.random_graph(4096, 40000, seed=5)[:2]
    csr = O.CSR(off, dst)
    full, st_full = O.cc_ref(csr, nthreads=1)
    with O.edge_limit(5000):
        _, st = O.cc_ref(csr, nthreads=1)
    maxdeg = int(np.diff(off).max())
    assert 5000 <= st["edges_relaxed"] < 5000 + maxdeg
    assert st["iterations"] == 1
    again, st2 = O.cc_ref(csr, nthreads=1)
    assert np.array_equal(again, full) and st2["edges_relaxed"] == st_full["edges_relaxed"]
    outdeg = np.diff(off).astype(np.uint32)
    O.set_num_threads(1)
    try:
        with O.edge_limit(7000):
            O.pr_norm(csr, outdeg, 0.85, 5)
        assert 7000 <= O.last_pulled() < 7000 + maxdeg
        O.pr_norm(csr, outdeg, 0.85, 2)
        assert O.last_pulled() == 2 * len(dst)
    finally:
        O.set_num_threads(os.cpu_count() or 1)
