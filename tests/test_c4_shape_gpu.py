"""GPU parity at the shape of the out-of-memory configuration: R-MAT with
edge factor 32, weights uniform in [1, 255], a device edge budget of one
third, BFS and SSSP from the highest-degree vertex, in every streaming mode
(CalculateOpt.cu:14-207, 383-545; GraphMeta.cuh:310-372).

Weight bytes up to 254 cross the host link in the split 5-byte record form
(a u32 id plus a byte w - wmin), forced by stream_pack 2 on small graphs and
chosen by the library itself (stream_pack 1) when the ids need more than 24
bits, as they do at scale 30: there 30 id bits and 8 weight bits do not fit
one 32-bit word.
"""
import sys

import numpy as np
import pytest

import ascetic_amd as A
from ascetic_amd import _abi
from oracle import oracle as O
from tests import graphs as G

pytestmark = pytest.mark.gpu

MODES = {"gather": _abi.STREAM_GATHER, "zerocopy": _abi.STREAM_ZEROCOPY, "hybrid": _abi.STREAM_HYBRID}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def rmat_ef32():
    h = A.rmat(14, 32, seed=30, weighted=True, wmin=1, wmax=255)
    off, ew = h.offsets().copy(), h.edge_words().copy()
    h.free()
    src = int(np.argmax(np.diff(off)))
    csr = O.CSR(off, ew[:, 0], ew[:, 1])
    assert ew[:, 1].max() >= 250 and ew[:, 1].min() == 1
    return off, ew, src, O.dijkstra(csr, src), O.bfs_queue(csr, src), O.dijkstra(csr, 7)


@pytest.mark.parametrize("pack,rec", [(2, 5), (1, 4), (0, 8)])
@pytest.mark.parametrize("mode", list(MODES))
def test_rmat_ef32_w255_budget_third(rmat_ef32, mode, pack, rec):
    off, ew, src, d_ref, b_ref, d7_ref = rmat_ef32
    E = int(off[-1])
    say("c4 shape", mode, "pack", pack)
    with A.Context(edge_budget_bytes=8 * E // 3, part_bytes=rec * 8192, stream_mode=MODES[mode],
                   stream_pack=pack) as ctx, ctx.from_arrays(off, ew) as g:
        info = g.info()
        assert 0.25 < info["static_edges"] / E <= 0.34
        d, st = g.sssp(src)
        assert np.array_equal(d, d_ref)
        assert st["streamed_edges"] > 0 and st["streamed_bytes"] == st["streamed_edges"] * rec
        lv, st = g.bfs(src)
        assert np.array_equal(lv, b_ref)
        assert st["streamed_bytes"] == st["streamed_edges"] * rec
        d, _ = g.sssp(7)        # a source inside the on-demand region's reach
        assert np.array_equal(d, d7_ref)


def _wide_id_graph():
    """A sparse symmetric graph over V = 2^24 + 4097 vertices: a random tree
    through 200k vertices spread over the whole id range (ids above 2^24
    included), plus 200k random chords among them; weights in [1, 255],
    equal in both directions."""
    V = (1 << 24) + 4097
    rng = np.random.default_rng(44)
    n = 200_000
    verts = rng.choice(V - 1, n - 1, replace=False)
    verts = np.concatenate([[V - 1], verts])         # the largest id is on the tree
    parent = verts[(rng.random(n - 1) * np.arange(1, n)).astype(np.int64)]
    u = np.concatenate([verts[1:], verts[rng.integers(0, n, n)]])
    v = np.concatenate([parent, verts[rng.integers(0, n, n)]])
    keep = u != v
    u, v = u[keep], v[keep]
    lo, hi = np.minimum(u, v), np.maximum(u, v)
    key = np.unique(lo.astype(np.int64) * V + hi)
    lo, hi = key // V, key % V
    w = (rng.integers(1, 256, len(lo))).astype(np.int64)
    w[0] = 255
    w[1] = 1
    su = np.concatenate([lo, hi])
    sv = np.concatenate([hi, lo])
    sw = np.concatenate([w, w])
    order = np.lexsort((sv, su))
    su, sv, sw = su[order], sv[order], sw[order]
    off = np.zeros(V + 1, np.uint64)
    off[1:] = np.cumsum(np.bincount(su, minlength=V))
    return off, sv.astype(np.uint32), sw.astype(np.uint32)


@pytest.mark.parametrize("mode", list(MODES))
def test_split_form_chosen_for_wide_ids(mode):
    """stream_pack 1 picks the 5-byte split form by itself once 25 id bits and
    8 weight bits exceed one word; values stay bit-exact."""
    off, dst, w = _wide_id_graph()
    V, E = len(off) - 1, int(off[-1])
    csr = O.CSR(off, dst, w)
    deg = np.diff(off)
    src = int(np.argmax(deg))
    far = int(dst[off[V - 1]])             # a neighbour of the largest id
    say("wide ids", mode, "V", V, "E", E)
    with A.Context(edge_budget_bytes=8 * E // 3, part_bytes=5 * 16384, stream_mode=MODES[mode],
                   stream_pack=1) as ctx, ctx.from_arrays(off, G.wedges(dst, w)) as g:
        assert g.info()["ondemand_edges"] > 0
        for s in (src, far):
            d, st = g.sssp(s)
            assert np.array_equal(d, O.dijkstra(csr, s)), s
            assert st["streamed_edges"] > 0 and st["streamed_bytes"] == 5 * st["streamed_edges"]
            lv, _ = g.bfs(s)
            assert np.array_equal(lv, O.bfs_queue(csr, s)), s


@pytest.mark.parametrize("pack", [1, 2])
def test_release_host_edges_after_packing(rmat_ef32, pack):
    """asc_config.release_host_edges frees the graph's host {dst, w} copy once
    the on-demand edges sit in their packed pinned form: the host then holds
    offsets and packed records only, results are unchanged, and re-planning
    (which would need the records) is refused."""
    off, ew, src, d_ref, b_ref, _ = rmat_ef32
    E = int(off[-1])
    say("release host edges, pack", pack)
    kw = dict(edge_budget_bytes=8 * E // 3, part_bytes=5 * 8192, stream_pack=pack)
    with A.Context(**kw) as c0, c0.from_arrays(off, ew) as g0:
        kept = g0.info()["host_bytes"]
    with A.Context(release_host_edges=True, **kw) as ctx, ctx.from_arrays(off, ew) as g:
        info = g.info()
        rec = 5 if pack == 2 else 4
        assert info["host_bytes"] == 8 * (len(off)) + rec * info["ondemand_edges"]
        assert info["host_bytes"] < kept - 8 * E // 2
        d, _ = g.sssp(src)
        assert np.array_equal(d, d_ref)
        lv, _ = g.bfs(src)
        assert np.array_equal(lv, b_ref)
        with pytest.raises(A.AscError) as e:
            g.replan(edge_budget_bytes=8 * E // 2)
        assert e.value.code == _abi.ERR_UNSUPPORTED


@pytest.mark.parametrize("pack,rec", [(2, 5), (1, 4)])
def test_pageable_tail(rmat_ef32, pack, rec):
    """A tail larger than max_pinned_bytes (as at scale 30, where the host
    grants about 70 GB of page-locked memory for 110 GB of records) lives in
    pageable memory: dense chunks go through pinned staging, results are the
    oracle's; zero-copy and hybrid streaming, which need page-locked records,
    are refused."""
    off, ew, src, d_ref, b_ref, d7_ref = rmat_ef32
    E = int(off[-1])
    say("c4 shape pageable tail, pack", pack)
    with A.Context(edge_budget_bytes=8 * E // 3, part_bytes=rec * 8192, stream_pack=pack,
                   max_pinned_bytes=1) as ctx, ctx.from_arrays(off, ew) as g:
        assert g.info()["tail_pinned"] == 0
        for _ in range(2):   # the second run reuses chunks left in the part buffers
            d, st = g.sssp(src)
            assert np.array_equal(d, d_ref)
            assert st["dense_edges"] > 0 and st["streamed_bytes"] == st["streamed_edges"] * rec
        lv, _ = g.bfs(src)
        assert np.array_equal(lv, b_ref)
        d, _ = g.sssp(7)
        assert np.array_equal(d, d7_ref)
    with A.Context(edge_budget_bytes=8 * E // 3, part_bytes=rec * 8192, stream_pack=pack,
                   max_pinned_bytes=0) as ctx, ctx.from_arrays(off, ew) as g:
        assert g.info()["tail_pinned"] == 1
    for mode in ("zerocopy", "hybrid"):
        with A.Context(edge_budget_bytes=8 * E // 3, part_bytes=rec * 8192, stream_pack=pack, max_pinned_bytes=1,
                       stream_mode=MODES[mode]) as ctx, ctx.from_arrays(off, ew) as g:
            with pytest.raises(A.AscError):
                g.sssp(src)
