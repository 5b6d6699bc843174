"""GPU serving: continuous batching of BFS / SSSP queries.

Each query's result in the mix must equal the same query run alone through
the single-query engine and the oracle, bit for bit."""
import random
import sys
import threading
import time

import numpy as np
import pytest

import ascetic_amd as A
from ascetic_amd import _abi
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def wgraph(gpu_ctx):
    h = A.rmat(15, 16, seed=11, weighted=True, wmin=1, wmax=64)
    off, ew = h.offsets().copy(), h.edge_words().copy()
    g = gpu_ctx.from_arrays(off, ew)
    yield g, off, ew
    g.close()


def test_mix_equals_alone(wgraph):
    g, off, ew = wgraph
    V = len(off) - 1
    deg = np.diff(off)
    rng = np.random.default_rng(3)
    cand = np.nonzero(deg > 0)[0]
    srcs = [int(x) for x in rng.choice(cand, 150, replace=False)]
    kinds = [A.Server.SSSP if rng.random() < 0.3 else A.Server.BFS for _ in srcs]
    csr = O.CSR(off, ew[:, 0], ew[:, 1])
    outs = [np.empty(V, np.uint32) for _ in srcs]
    say("serving 150 queries")
    with A.Server(g, lanes=64) as srv:
        tickets = [srv.submit(k, s, o) for k, s, o in zip(kinds, srcs, outs)]
        stats = [srv.wait(t) for t in tickets]
        st = srv.stats()
    assert st["completed"] == 150 and st["lanes_bfs"] == 64 and st["lanes_sssp"] == 64
    assert max(s["lane"] for s in stats) > 0
    for k, s, o in zip(kinds, srcs, outs):
        alone, _ = (g.sssp(s) if k == A.Server.SSSP else g.bfs(s))
        assert np.array_equal(o, alone), (k, s)
    # spot-check against the oracle
    for k, s, o in list(zip(kinds, srcs, outs))[:12]:
        ref = O.dijkstra(csr, s) if k == A.Server.SSSP else O.bfs_queue(csr, s)
        assert np.array_equal(o, ref), (k, s)


def test_step_graphs_replay_exactly(wgraph):
    """Engine steps replayed from recorded hipGraphs (the default) give the
    same results as steps launched kernel by kernel: many replays of both
    parities, with admissions (eager seeding kernels) and host syncs between
    them, every result compared."""
    g, off, ew = wgraph
    V = len(off) - 1
    deg = np.diff(off)
    rng = np.random.default_rng(13)
    cand = np.nonzero(deg > 0)[0]
    srcs = [int(x) for x in rng.choice(cand, 120, replace=False)]
    kinds = [A.Server.SSSP if i % 3 == 0 else A.Server.BFS for i in range(len(srcs))]
    res = {}
    for graphs in (True, False):
        outs = [np.empty(V, np.uint32) for _ in srcs]
        say("serving with step graphs", graphs)
        with A.Server(g, lanes=32, step_graphs=graphs) as srv:
            # two waves, so that lanes are admitted while others are in flight
            tickets = [srv.submit(k, s, o) for k, s, o in zip(kinds[:60], srcs[:60], outs[:60])]
            for t in tickets[:20]:
                srv.wait(t)
            tickets += [srv.submit(k, s, o) for k, s, o in zip(kinds[60:], srcs[60:], outs[60:])]
            for t in tickets[20:]:
                srv.wait(t)
            st = srv.stats()
        if graphs:
            assert st["graph_steps"] >= 8 and st["graph_steps"] >= st["steps"] // 2
        else:
            assert st["graph_steps"] == 0
        res[graphs] = outs
    for i, (k, s) in enumerate(zip(kinds, srcs)):
        assert np.array_equal(res[True][i], res[False][i]), (k, s)
    for i, (k, s) in enumerate(list(zip(kinds, srcs))[:30]):
        alone, _ = (g.sssp(s) if k == A.Server.SSSP else g.bfs(s))
        assert np.array_equal(res[True][i], alone), (k, s)


def test_bottom_up_lanes_exact(wgraph):
    """BFS lanes that switch to bottom-up steps (symmetric graph) give the
    same levels as top-down-only serving, single runs and the oracle."""
    g, off, ew = wgraph
    V = len(off) - 1
    deg = np.diff(off)
    rng = np.random.default_rng(7)
    cand = np.nonzero(deg > 0)[0]
    srcs = [int(x) for x in rng.choice(cand, 100, replace=False)] + [0, int(np.argmax(deg))]
    csr = O.CSR(off, ew[:, 0], ew[:, 1])
    res = {}
    for direction in (_abi.BFS_AUTO, _abi.BFS_TOP_DOWN):
        outs = [np.empty(V, np.uint32) for _ in srcs]
        say("bfs serving, direction", direction)
        with A.Server(g, lanes=64, bfs_direction=direction) as srv:
            tickets = [srv.submit(A.Server.BFS, s, o) for s, o in zip(srcs, outs)]
            for t in tickets:
                srv.wait(t)
            st = srv.stats()
        if direction == _abi.BFS_AUTO:
            assert st["bottom_up_steps"] > 0
        else:
            assert st["bottom_up_steps"] == 0
        res[direction] = outs
    for i, s in enumerate(srcs):
        assert np.array_equal(res[_abi.BFS_AUTO][i], res[_abi.BFS_TOP_DOWN][i]), s
        alone, _ = g.bfs(s)
        assert np.array_equal(res[_abi.BFS_AUTO][i], alone), s
    for i, s in enumerate(srcs[:10]):
        assert np.array_equal(res[_abi.BFS_AUTO][i], O.bfs_queue(csr, s)), s


def test_sssp_lanes_beyond_16_bits_fall_back_exactly():
    """Distances above 65534 do not fit a 16-bit lane: such queries are
    answered by the single-query engine, and every result stays exact."""
    from tests import graphs as G
    n = 40
    off, dst, _ = G.path_graph(n)
    w = np.full(len(dst), 30000, np.uint32)
    w[::3] = 7
    off2, dst2, w2 = G.random_graph(200, 600, 21, symmetric=True, weighted=True, wmax=64)
    # two components: a heavy path (overflows) and a light random graph
    offs = np.concatenate([off, off2[1:] + off[-1]]).astype(np.uint64)
    dsts = np.concatenate([dst, dst2 + n]).astype(np.uint32)
    ws = np.concatenate([w, w2]).astype(np.uint32)
    V = len(offs) - 1
    csr = O.CSR(offs, dsts, ws)
    with A.Context() as ctx, ctx.from_arrays(offs, np.stack([dsts, ws], axis=1)) as g:
        srcs = [0, 5, n - 1, n, n + 17, n + 150]
        outs = [np.empty(V, np.uint32) for _ in srcs]
        with A.Server(g, lanes=8) as srv:
            ts = [srv.submit(A.Server.SSSP, s_, o) for s_, o in zip(srcs, outs)]
            for t in ts:
                srv.wait(t)
            st = srv.stats()
        assert st["fallbacks"] >= 3
        for s_, o in zip(srcs, outs):
            assert np.array_equal(o, O.dijkstra(csr, s_)), s_


def test_bfs_lanes_beyond_8_bits_fall_back_exactly():
    """Levels above 254 do not fit an 8-bit lane: such queries are answered
    by the single-query engine, and every result stays exact."""
    from tests import graphs as G
    n = 600
    off, dst, _ = G.path_graph(n)
    off2, dst2, _ = G.random_graph(200, 600, 23, symmetric=True)
    offs = np.concatenate([off, off2[1:] + off[-1]]).astype(np.uint64)
    dsts = np.concatenate([dst, dst2 + n]).astype(np.uint32)
    V = len(offs) - 1
    csr = O.CSR(offs, dsts)
    with A.Context() as ctx, ctx.from_arrays(offs, dsts) as g:
        srcs = [0, 100, n - 1, 300, n, n + 17, n + 150]
        outs = [np.empty(V, np.uint32) for _ in srcs]
        with A.Server(g, lanes=8) as srv:
            ts = [srv.submit(A.Server.BFS, s_, o) for s_, o in zip(srcs, outs)]
            for t in ts:
                srv.wait(t)
            st = srv.stats()
        assert st["fallbacks"] >= 3
        for s_, o in zip(srcs, outs):
            assert np.array_equal(o, O.bfs_queue(csr, s_)), s_


def test_served_values_follow_value_convention():
    h = A.rmat(12, 8, seed=4, weighted=True)
    off, ew = h.offsets().copy(), h.edge_words().copy()
    V = len(off) - 1
    with A.Context(value_convention=_abi.VALUES_ZERO_BASED) as ctx, ctx.from_arrays(off, ew) as g:
        outs = {k: np.empty(V, np.uint32) for k in (A.Server.BFS, A.Server.SSSP)}
        with A.Server(g, lanes=4) as srv:
            ts = [srv.submit(k, 3, o) for k, o in outs.items()]
            for t in ts:
                srv.wait(t)
        assert np.array_equal(outs[A.Server.BFS], g.bfs(3)[0])
        assert np.array_equal(outs[A.Server.SSSP], g.sssp(3)[0])
        assert outs[A.Server.BFS][3] == 0 and outs[A.Server.SSSP][3] == 0


def test_staggered_arrivals_from_threads(wgraph):
    g, off, ew = wgraph
    V = len(off) - 1
    cand = np.nonzero(np.diff(off) > 0)[0]
    results = {}
    errors = []
    with A.Server(g, lanes=16) as srv:
        def client(seed):
            r = random.Random(seed)
            try:
                for _ in range(12):
                    s = int(cand[r.randrange(len(cand))])
                    k = A.Server.SSSP if r.random() < 0.3 else A.Server.BFS
                    out = np.empty(V, np.uint32)
                    t = srv.submit(k, s, out)
                    time.sleep(r.random() * 0.004)
                    srv.wait(t)
                    results[(seed, _)] = (k, s, out)
            except Exception as e:  # pragma: no cover - reported below
                errors.append(e)

        threads = [threading.Thread(target=client, args=(i,)) for i in range(6)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
    assert not errors, errors
    assert len(results) == 72
    for k, s, out in results.values():
        alone, _ = (g.sssp(s) if k == A.Server.SSSP else g.bfs(s))
        assert np.array_equal(out, alone), (k, s)


def test_lanes_limited_by_memory_budget(wgraph):
    g, off, ew = wgraph
    V = len(off) - 1
    budget = 2 * 3 * V  # two lanes per kind: V bytes per BFS lane, 2 V per SSSP lane
    srcs = [0, 5, 17, 99, 1000, 2047]
    outs = [np.empty(V, np.uint32) for _ in srcs]
    with A.Server(g, state_budget_bytes=budget) as srv:
        assert srv.stats()["lanes_bfs"] == 2
        ts = [srv.submit(A.Server.SSSP if i % 2 else A.Server.BFS, s, o) for i, (s, o) in enumerate(zip(srcs, outs))]
        for t in ts:
            srv.wait(t)
    for i, (s, o) in enumerate(zip(srcs, outs)):
        alone, _ = g.sssp(s) if i % 2 else g.bfs(s)
        assert np.array_equal(o, alone)


def test_serving_errors(gpu_ctx):
    h = A.rmat(10, 8, seed=2)
    with gpu_ctx.adopt(h) as g:
        with A.Server(g, queue_capacity=2) as srv:
            with pytest.raises(A.AscError) as e:
                srv.submit(A.Server.SSSP, 0)
            assert e.value.code == _abi.ERR_UNSUPPORTED
            with pytest.raises(A.AscError) as e:
                srv.submit(A.Server.BFS, 1 << 20)
            assert e.value.code == _abi.ERR_INVALID
            # result buffers the library could overrun are refused before submission
            # (adopt() took the host graph over, so V comes from the device graph)
            V = g.num_vertices
            for bad in (np.empty(V - 1, np.uint32), np.empty(V, np.int64), np.empty(2 * V, np.uint32)[::2]):
                with pytest.raises(ValueError):
                    srv.submit(A.Server.BFS, 0, bad)
            srv.drain()
