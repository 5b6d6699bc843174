"""asc_graph_borrow_csr: a graph over the caller's arrays without copying
them gives the oracle's values resident and streamed (the gather then reads
the caller's records), with u64 offsets borrowed and u32 ones widened, and
the library's host footprint excludes the borrowed arrays."""
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
def weighted():
    off, dst, w = G.random_graph(20000, 160000, 51, symmetric=True, weighted=True)
    ew = G.wedges(dst, w)
    csr = O.CSR(off, dst, w)
    return off.astype(np.uint64), ew, csr


@pytest.mark.parametrize("pack", [0, 1])
def test_borrowed_graph_streamed(weighted, pack):
    off, ew, csr = weighted
    E = int(off[-1])
    say("borrow streamed, pack", pack)
    keep = ew.copy()
    with A.Context(edge_budget_bytes=8 * E // 3, part_bytes=8 * 4096, stream_pack=pack) as ctx:
        g = ctx.borrow_arrays(off, ew)
        info = g.info()
        # borrowed offsets and records are not the library's memory: it holds
        # only the pinned on-demand copy (packed 4-byte words, or 8-byte
        # records for whole-chunk copies)
        assert info["host_bytes"] == (4 if pack else 8) * (E - info["static_edges"])
        for s in (0, 77):
            d, st = g.sssp(s)
            assert np.array_equal(d, O.dijkstra(csr, s)), s
            assert st["streamed_edges"] > 0
            lv, _ = g.bfs(s)
            assert np.array_equal(lv, O.bfs_queue(csr, s)), s
        g.close()
    assert np.array_equal(ew, keep)   # the caller's array is untouched


def test_borrowed_graph_u32_offsets_resident(weighted):
    off, ew, csr = weighted
    say("borrow u32 offsets")
    with A.Context() as ctx:
        g = ctx.borrow_arrays(off.astype(np.uint32), ew)
        info = g.info()
        assert info["host_bytes"] == 8 * (len(off))   # the widened offsets only
        d, _ = g.sssp(3)
        assert np.array_equal(d, O.dijkstra(csr, 3))
        lab, _ = g.cc()
        assert np.array_equal(lab, O.cc_unionfind(csr))
        g.close()
