"""Object lifetimes through the C ABI: re-planning a graph (a plan that cannot
fit is refused before anything changes; the budget belongs to the graph, not
its context), graphs with a live server, and one waiter per served ticket."""
import threading

import numpy as np
import pytest

import ascetic_amd as A
from ascetic_amd import _abi
from oracle import oracle as O
from tests import graphs as G

pytestmark = pytest.mark.gpu


def _graph():
    off, dst, _ = G.random_graph(20000, 200000, 13, symmetric=True)
    w = np.random.default_rng(13).integers(1, 65, len(dst)).astype(np.uint32)
    return off, dst, w


def test_replan_that_cannot_fit_keeps_the_graph():
    off, dst, w = _graph()
    V, E = len(off) - 1, int(off[-1])
    csr = O.CSR(off, dst, w)
    ref = O.dijkstra(csr, 0)
    limit = 8 * (V + 1) + 8 * E * 3 // 5          # offsets + 60% of the edge records
    with A.Context(device_mem_limit=limit, edge_budget_bytes=8 * E // 4) as ctx:
        with ctx.from_arrays(off, G.wedges(dst, w)) as g:
            before = g.info()
            assert 0 < before["static_edges"] < E
            # every edge resident would exceed the limit: refused, nothing changed
            with pytest.raises(A.AscError) as e:
                g.replan(edge_budget_bytes=8 * E)
            assert e.value.code == _abi.ERR_NOMEM
            assert g.info() == before
            d, _ = g.sssp(0)
            assert np.array_equal(d, ref)
            # a plan that fits replaces the split, for this graph only
            g.replan(edge_budget_bytes=8 * E // 2)
            after = g.info()
            assert before["static_edges"] < after["static_edges"] <= E // 2
            d, _ = g.sssp(0)
            assert np.array_equal(d, ref)
            with ctx.from_arrays(off, G.wedges(dst, w)) as g2:
                assert g2.info()["static_edges"] == before["static_edges"]


def test_sssp_across_replans_resident_and_streamed():
    """SSSP on one graph streamed, re-planned fully resident (where it probes
    a 16-bit value filter), then streamed again: the oracle's values each
    time, twice per plan (the traversal state is reused between runs)."""
    off, dst, w = _graph()
    E = int(off[-1])
    csr = O.CSR(off, dst, w)
    refs = {s: O.dijkstra(csr, s) for s in (0, 3)}
    with A.Context(edge_budget_bytes=8 * E // 3) as ctx, ctx.from_arrays(off, G.wedges(dst, w)) as g:
        for budget in (8 * E // 3, 8 * E, 8 * E // 4):
            g.replan(edge_budget_bytes=budget)
            resident = g.info()["static_edges"] == E
            assert resident == (budget == 8 * E)
            for s, ref in refs.items():
                d, st = g.sssp(s)
                assert np.array_equal(d, ref), (budget, s)
                assert (st["streamed_edges"] == 0) or not resident


def test_replan_and_destroy_refused_while_a_server_runs():
    off, dst, w = _graph()
    csr = O.CSR(off, dst, w)
    with A.Context() as ctx:
        g = ctx.from_arrays(off, G.wedges(dst, w))
        srv = A.Server(g, lanes=4)
        with pytest.raises(A.AscError) as e:
            g.replan(edge_budget_bytes=1 << 20)
        assert e.value.code == _abi.ERR_STATE
        with pytest.raises(A.AscError) as e:
            g.close()
        assert e.value.code == _abi.ERR_STATE
        out = np.empty(len(off) - 1, np.uint32)
        srv.wait(srv.submit(A.Server.SSSP, 3, out))      # the server still works
        assert np.array_equal(out, O.dijkstra(csr, 3))
        srv.close()
        g.replan(edge_budget_bytes=8 * int(off[-1]) // 3)
        d, _ = g.sssp(3)
        assert np.array_equal(d, O.dijkstra(csr, 3))
        g.close()


def test_one_waiter_per_ticket():
    """Two threads waiting on one ticket: one gets the result, the other
    ERR_INVALID (it never touches the freed query)."""
    off, dst, w = _graph()
    with A.Context() as ctx, ctx.from_arrays(off, G.wedges(dst, w)) as g, A.Server(g, lanes=8) as srv:
        for _ in range(5):
            t = srv.submit(A.Server.SSSP, 0)
            codes = []

            def waiter():
                try:
                    srv.wait(t)
                    codes.append(0)
                except A.AscError as e:
                    codes.append(e.code)

            th = [threading.Thread(target=waiter) for _ in range(2)]
            for x in th:
                x.start()
            for x in th:
                x.join()
            assert sorted(codes) == [_abi.ERR_INVALID, 0]
        with pytest.raises(A.AscError) as e:
            srv.wait(t)                                    # released
        assert e.value.code == _abi.ERR_INVALID
