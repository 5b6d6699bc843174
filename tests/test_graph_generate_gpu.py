"""asc_graph_generate: a weighted graph generated straight into the device's
static prefix and the pinned packed tail gives the oracle's results on the
host generator's CSR for the same parameters (BFS levels, SSSP distances,
CC labels, bit for bit), with the same edge count and plan, in every
streaming mode and both packed record forms, fully resident too.  The
budget-capped sizing follows GraphMeta.cuh:310-372; the traversals
CalculateOpt.cu:14-207, 383-545."""
import sys

import numpy as np
import pytest

import ascetic_amd as A
from ascetic_amd import _abi
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MODES = {"gather": _abi.STREAM_GATHER, "zerocopy": _abi.STREAM_ZEROCOPY, "hybrid": _abi.STREAM_HYBRID}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


PARAMS = {
    "rmat14_ef32_w255": dict(type=0, scale=14, edge_factor=32, seed=30, wmin=1, wmax=255),
    "er_w64": dict(type=1, num_vertices=50021, num_samples=400000, seed=8, wmin=1, wmax=64),
}


@pytest.fixture(scope="module")
def reference():
    out = {}
    for name, kw in PARAMS.items():
        h = A.generate(weighted=True, **kw)
        off, ew = h.offsets().copy(), h.edge_words().copy()
        h.free()
        csr = O.CSR(off, ew[:, 0], ew[:, 1])
        src = int(np.argmax(np.diff(off)))
        out[name] = (off, ew, src, O.bfs_queue(csr, src), O.dijkstra(csr, src), O.cc_unionfind(csr))
    return out


@pytest.mark.parametrize("pack,rec", [(1, 4), (2, 5)])
@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("name", list(PARAMS))
def test_generated_graph_matches_oracle(reference, name, mode, pack, rec):
    off, ew, src, b_ref, d_ref, c_ref = reference[name]
    E = int(off[-1])
    say("graph-generate", name, mode, "pack", pack)
    with A.Context(edge_budget_bytes=8 * E // 3, part_bytes=rec * 8192, stream_mode=MODES[mode],
                   stream_pack=pack) as ctx, ctx.generate_graph(**PARAMS[name]) as g, \
            A.Context(edge_budget_bytes=8 * E // 3, stream_pack=pack) as c2, c2.from_arrays(off, ew) as g2:
        info, info2 = g.info(), g2.info()
        assert info["num_edges"] == E and info["num_vertices"] == len(off) - 1
        assert info["static_edges"] == info2["static_edges"] and info["static_vertices"] == info2["static_vertices"]
        assert 0.25 < info["static_edges"] / E <= 0.34
        lv, st = g.bfs(src)
        assert np.array_equal(lv, b_ref)
        assert st["streamed_edges"] > 0 and st["streamed_bytes"] == st["streamed_edges"] * rec
        d, st = g.sssp(src)
        assert np.array_equal(d, d_ref)
        d2, st2 = g2.sssp(src)
        assert np.array_equal(d2, d_ref)
        lab, _ = g.cc()
        assert np.array_equal(lab, c_ref)


def test_generated_graph_resident(reference):
    """Everything on the device: the symmetry check passes and the resident
    fast paths (direction-optimizing BFS, union-find CC) give the oracle's
    values."""
    off, ew, src, b_ref, d_ref, c_ref = reference["rmat14_ef32_w255"]
    say("graph-generate resident")
    with A.Context() as ctx, ctx.generate_graph(**PARAMS["rmat14_ef32_w255"]) as g:
        info = g.info()
        assert info["static_edges"] == int(off[-1])
        lv, st = g.bfs(src)
        assert np.array_equal(lv, b_ref) and st["path"] == "direction_opt"
        d, _ = g.sssp(src)
        assert np.array_equal(d, d_ref)
        lab, st = g.cc()
        assert np.array_equal(lab, c_ref) and st["path"] == "union_find"


def test_generated_graph_refuses_replan_and_unpacked():
    """A generated graph has no host records to re-plan from, and its
    on-demand edges exist only in packed form."""
    with A.Context(edge_budget_bytes=1 << 18) as ctx, \
            ctx.generate_graph(type=0, scale=12, edge_factor=16, seed=2, wmax=64) as g:
        assert 0 < g.info()["static_edges"] < g.info()["num_edges"]
        with pytest.raises(A.AscError):
            g.replan(edge_budget_bytes=1 << 22)
    with A.Context(edge_budget_bytes=1 << 18, stream_pack=0) as ctx:
        with pytest.raises(A.AscError):
            ctx.generate_graph(type=0, scale=12, edge_factor=16, seed=2, wmax=64)


def test_generated_graph_pageable_tail(reference):
    """The generator's tail in pageable memory (max_pinned_bytes below its
    size): gather streaming through pinned staging gives the oracle's values."""
    off, ew, src, b_ref, d_ref, c_ref = reference["rmat14_ef32_w255"]
    E = int(off[-1])
    say("graph-generate pageable tail")
    with A.Context(edge_budget_bytes=8 * E // 3, part_bytes=5 * 8192, stream_pack=2, max_pinned_bytes=1) as ctx, \
            ctx.generate_graph(**PARAMS["rmat14_ef32_w255"]) as g:
        assert g.info()["tail_pinned"] == 0
        lv, _ = g.bfs(src)
        assert np.array_equal(lv, b_ref)
        d, st = g.sssp(src)
        assert np.array_equal(d, d_ref) and st["dense_edges"] > 0
