"""GPU autotuner: the sweep finishes within its budget, never picks values
slower than the defaults on the sampled workload, changes no result, and
round-trips through a tuning table keyed by the graph."""
import numpy as np
import pytest

import ascetic_amd as A
from ascetic_amd import _abi
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_autotune_traversal_graph(gpu_ctx, tmp_path):
    h = A.rmat(16, 16, seed=5, weighted=True)
    off, ew = h.offsets().copy(), h.edge_words().copy()
    csr = O.CSR(off, ew[:, 0], ew[:, 1])
    src = int(np.argmax(np.diff(off)))
    with gpu_ctx.from_arrays(off, ew) as g:
        t = g.autotune(budget_ms=20000)
        assert t["score_ms"] <= t["untuned_ms"]
        assert t["heavy_threshold"] >= 256 and t["blocks_per_cu"] in (2, 4, 8, 16)
        assert t["serve_lanes"] in (16, 32, 64)
        d, _ = g.sssp(src)
        assert np.array_equal(d, O.dijkstra(csr, src))
        lv, _ = g.bfs(src)
        assert np.array_equal(lv, O.bfs_queue(csr, src))
        path = str(tmp_path / "tuning.txt")
        g.save_tuning(path)
        g.save_tuning(path)  # replaces its own line
        assert len(open(path).read().strip().splitlines()) == 1
    with gpu_ctx.from_arrays(off, ew) as g2:
        t2 = g2.load_tuning(path)
        assert t2["heavy_threshold"] == t["heavy_threshold"] and t2["serve_lanes"] == t["serve_lanes"]
    other = A.rmat(12, 8, seed=6)
    with gpu_ctx.adopt(other) as g3:
        with pytest.raises(A.AscError) as e:
            g3.load_tuning(path)
        assert e.value.code == _abi.ERR_STATE


def test_autotune_pagerank_graph(gpu_ctx):
    h = A.rmat(14, 16, seed=7, csc=True)
    ioff, isrc, od = h.offsets().copy(), h.dst().copy(), h.out_degree().copy()
    with gpu_ctx.adopt(h) as g:
        t = g.autotune(budget_ms=10000)
        assert t["pr_group"] in (0, 4, 8, 16, 32, 64)
        r, _, _ = g.pagerank(iterations=20)
    ref, _ = O.pr_norm(O.CSR(ioff, isrc), od, 0.85, 20)
    assert np.abs(r.astype(np.float64) - ref).sum() <= 1e-5
