"""Hand-derived known answers (tests/golden/known_answers.json): the oracle on
the CPU, and the library on the GPU through the C ABI, with the graph fully
resident and with every edge on demand (streamed from host memory).  The
cases include a single vertex, graphs without edges, self-loops, parallel
edges of different weights and a source without out-edges."""
import json
import os

import numpy as np
import pytest

import ascetic_amd as A
from ascetic_amd import _abi
from oracle import oracle as O
from tests import graphs as G

with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "known_answers.json")) as f:
    CASES = json.load(f)["cases"]


def build(case):
    e = case["edges"]
    weighted = case.get("weighted", bool(e) and len(e[0]) == 3)
    off, dst, w = G.csr_from_edges(case["V"], [(x[0], x[1]) for x in e], [x[2] for x in e] if weighted else None,
                                   symmetric=case["symmetric"], dedup=case.get("dedup", True))
    return off, dst, w


def oracle_answer(case):
    off, dst, w = build(case)
    V = case["V"]
    if case["algo"] == "pagerank":
        ioff, isrc = G.transpose(off, dst, V)
        r, _ = O.pr_norm(O.CSR(ioff, isrc), np.diff(off).astype(np.uint32), case["damping"], case["iterations"])
        return r
    csr = O.CSR(off, dst, w)
    if case["algo"] == "bfs":
        out = O.bfs_queue(csr, case["source"])
    elif case["algo"] == "sssp":
        out = O.dijkstra(csr, case["source"])
    else:
        out = O.cc_minreach(csr)
    if case.get("convention") == "zero":
        out = np.where(out == V + 1, 0xFFFFFFFF, out - 1).astype(np.uint32)
    return out


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_matches_known_answer(case):
    got = oracle_answer(case)
    if case["algo"] == "pagerank":
        assert np.allclose(got, case["expected"], rtol=0, atol=1e-12)
    else:
        assert got.tolist() == case["expected"]


def test_huge_weight_value_is_the_saturating_one():
    """Which oracle function defines SSSP when a value plus a weight leaves
    32 bits.  The reference adds in plain 32-bit `unsigned` arithmetic
    (gpu_kernels.cuh:496-499), and so does the oracle's restatement of it,
    orc_sssp_ref: 1 + (2^32 - 1) wraps to 0, which is accepted as smaller than
    V + 1, and the restatement answers distances below the true ones (an
    overflow of the reference, not a graph property).  The library adds with
    saturation instead, so its value is the exact distance capped at V + 1:
    orc_dijkstra's (64-bit sums, min(1 + d, V + 1)).  The golden case
    huge_weight3_sssp_from_0 records that value, and the GPU tests above hold
    the library to it; below 2^32 - V - 1 the two additions agree."""
    case = next(c for c in CASES if c["name"] == "huge_weight3_sssp_from_0")
    off, dst, w = build(case)
    csr = O.CSR(off, dst, w)
    assert O.dijkstra(csr, 0).tolist() == case["expected"] == [1, 4, 4]
    wrapped, _ = O.sssp_ref(csr, 0)
    assert wrapped.tolist() == [1, 0, 1]          # 1 + (2^32 - 1) = 0 (mod 2^32), then 0 + 1
    # one below the wrap point the restatement and Dijkstra agree
    e = [[0, 1, 2 ** 32 - 1 - 4], [1, 2, 1]]
    off, dst, w = G.csr_from_edges(3, [(a, b) for a, b, _ in e], [c for _, _, c in e])
    csr = O.CSR(off, dst, w)
    assert O.sssp_ref(csr, 0)[0].tolist() == O.dijkstra(csr, 0).tolist() == [1, 4, 4]


@pytest.mark.gpu
@pytest.mark.parametrize("on_demand", [False, True], ids=["resident", "on_demand"])
@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_library_matches_known_answer(case, on_demand):
    off, dst, w = build(case)
    V = case["V"]
    if on_demand and case["algo"] == "pagerank":
        pytest.skip("partially resident PageRank is covered by test_pagerank_gpu")
    conv = _abi.VALUES_ZERO_BASED if case.get("convention") == "zero" else _abi.VALUES_REFERENCE
    # a one-byte device edge budget leaves every edge on demand
    budget = dict(edge_budget_bytes=1, part_bytes=4096) if on_demand else {}
    print("case", case["name"], "on_demand", on_demand, flush=True)
    with A.Context(value_convention=conv, **budget) as ctx:
        if case["algo"] == "pagerank":
            ioff, isrc = G.transpose(off, dst, V)
            with ctx.from_arrays(ioff, isrc, out_degree=np.diff(off).astype(np.uint32)) as g:
                r, _, _ = g.pagerank(iterations=case["iterations"], damping=case["damping"])
            assert float(np.abs(r.astype(np.float64) - np.array(case["expected"])).sum()) <= 1e-5
            return
        edges = dst if w is None else G.wedges(dst, w)
        with ctx.from_arrays(off, edges) as g:
            if case["algo"] == "bfs":
                out, _ = g.bfs(case["source"])
            elif case["algo"] == "sssp":
                out, _ = g.sssp(case["source"])
            else:
                out, _ = g.cc()
        assert out.tolist() == case["expected"]


SERVED = [c for c in CASES if c["algo"] in ("bfs", "sssp")]


@pytest.mark.gpu
@pytest.mark.parametrize("step_graphs", [True, False], ids=["graphs", "eager"])
@pytest.mark.parametrize("case", SERVED, ids=[c["name"] for c in SERVED])
def test_served_matches_known_answer(case, step_graphs):
    """The serving engine answers each case exactly, alone and mixed with
    queries from every vertex (several copies of the case's query among them),
    including the cases whose values reach the V+1 cap."""
    off, dst, w = build(case)
    V = case["V"]
    conv = _abi.VALUES_ZERO_BASED if case.get("convention") == "zero" else _abi.VALUES_REFERENCE
    kind = A.Server.BFS if case["algo"] == "bfs" else A.Server.SSSP
    print("served case", case["name"], "step_graphs", step_graphs, flush=True)
    with A.Context(value_convention=conv) as ctx:
        edges = dst if w is None else G.wedges(dst, w)
        with ctx.from_arrays(off, edges) as g:
            with A.Server(g, lanes=16, step_graphs=step_graphs) as srv:
                alone = np.empty(V, np.uint32)
                srv.wait(srv.submit(kind, case["source"], alone))
                assert alone.tolist() == case["expected"]
                mix = [case["source"]] * 3 + list(range(V))
                outs = [np.empty(V, np.uint32) for _ in mix]
                tickets = [srv.submit(kind, s, o) for s, o in zip(mix, outs)]
                for t in tickets:
                    srv.wait(t)
            for s, o in zip(mix, outs):
                ref, _ = (g.bfs(s) if kind == A.Server.BFS else g.sssp(s))
                assert np.array_equal(o, ref), (case["name"], s)
            for o in outs[:3]:
                assert o.tolist() == case["expected"]
