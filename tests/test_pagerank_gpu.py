"""GPU parity for PageRank through the C ABI.

Normalised mode (FP32 ranks, FP64 residual): L1 error <= 1e-5 against the
FP64 oracle after the same number of iterations.  Reference mode (FP64
active set, gpu_kernels.cuh:533-594): sums run sequentially in CSR order
with no fused multiply-add, as in the oracle, so ranks and iteration counts
must be identical.
"""
import numpy as np
import pytest

import ascetic_amd as A
from ascetic_amd import _abi
from oracle import oracle as O
from tests import graphs as G

pytestmark = pytest.mark.gpu

TOL_L1 = 1e-5


def csc_of(off, dst, V):
    ioff, isrc = G.transpose(off, dst, V)
    return ioff, isrc, np.diff(off).astype(np.uint32)


@pytest.mark.parametrize("symmetric", [True, False])
@pytest.mark.parametrize("V,M,seed", [(1000, 5000, 1), (20000, 200000, 2)])
def test_pr_normalised_l1(gpu_ctx, symmetric, V, M, seed):
    off, dst, _ = G.random_graph(V, M, seed, symmetric=symmetric)
    ioff, isrc, od = csc_of(off, dst, V)
    ref, ref_res = O.pr_norm(O.CSR(ioff, isrc), od, 0.85, 20)
    with gpu_ctx.from_arrays(ioff, isrc, out_degree=od) as g:
        r, res, st = g.pagerank(iterations=20)
    assert np.abs(r.astype(np.float64) - ref).sum() <= TOL_L1
    assert np.allclose(res, ref_res, rtol=1e-3, atol=1e-7)
    assert st["iterations"] == 20


def test_pr_normalised_hub_and_dangling(gpu_ctx):
    # a hub with in-degree above the workgroup threshold and dangling vertices
    V = 9000
    # vertices 5000.. have no out-edges; vertex 0 has in-degree 4999
    edges = [(i, 0) for i in range(1, 5000)] + [(0, i) for i in range(1, 50)] + [(i, i + 1) for i in range(1, 4000)]
    off, dst, _ = G.csr_from_edges(V, edges)
    ioff, isrc, od = csc_of(off, dst, V)
    assert (od == 0).any()
    ref, _ = O.pr_norm(O.CSR(ioff, isrc), od, 0.85, 20)
    with gpu_ctx.from_arrays(ioff, isrc, out_degree=od) as g:
        r, _, _ = g.pagerank(iterations=20)
    assert np.abs(r.astype(np.float64) - ref).sum() <= TOL_L1


@pytest.mark.parametrize("case", ["sym", "dir", "hub", "rmat", "multi_bin", "multi_segment"])
@pytest.mark.parametrize("method", [_abi.PR_PULL, _abi.PR_BLOCKED, _abi.PR_TILED])
def test_pr_blocked_and_pull_agree(gpu_ctx, case, method):
    """Propagation-blocked and pulled normalised PageRank both meet the L1
    bound: symmetric graphs (out-CSR = in-CSR), directed ones (transposed on
    the device), heavy destinations (kept on the FP64 pull), R-MAT and a
    graph spanning several 32768-vertex bins."""
    if case in ("sym", "dir"):
        V = 3000
        off, dst, _ = G.random_graph(V, 20000, 31, symmetric=case == "sym")
    elif case == "hub":
        V = 9000
        edges = [(i, 0) for i in range(1, 6000)] + [(0, i) for i in range(1, 70)] + [(i, i + 1) for i in range(1, 4000)]
        off, dst, _ = G.csr_from_edges(V, edges)
    elif case == "rmat":
        h = A.rmat(13, 16, seed=8)
        off, dst = h.offsets().copy(), h.dst().copy()
        V = len(off) - 1
    elif case == "multi_bin":
        V = 100000
        off, dst, _ = G.random_graph(V, 400000, 32, symmetric=False)
    else:   # three source segments of the tiled schedule, 37 destination blocks
        V = 600000
        off, dst, _ = G.random_graph(V, 1500000, 34, symmetric=False)
    ioff, isrc, od = csc_of(off, dst, V)
    ref, ref_res = O.pr_norm(O.CSR(ioff, isrc), od, 0.85, 20)
    with gpu_ctx.from_arrays(ioff, isrc, out_degree=od) as g:
        r, res, st = g.pagerank(iterations=20, method=method)
        r2, _, _ = g.pagerank(iterations=20, method=method)
    assert np.abs(r.astype(np.float64) - ref).sum() <= TOL_L1, case
    assert np.abs(r2.astype(np.float64) - ref).sum() <= TOL_L1, case
    assert np.allclose(res, ref_res, rtol=1e-3, atol=1e-7)


@pytest.mark.parametrize("sb,db", [(15, 14), (15, 15), (12, 15), (13, 13), (17, 15)])
def test_pr_tile_shapes(gpu_ctx, sb, db):
    """Every source-segment / destination-block shape of the tiled pull meets
    the L1 bound, 2^15-destination blocks (128 KB of LDS sums, one workgroup
    per CU) included, on a directed graph with a ragged last block."""
    V = 600000
    off, dst, _ = G.random_graph(V, 1500000, 35, symmetric=False)
    ioff, isrc, od = csc_of(off, dst, V)
    ref, _ = O.pr_norm(O.CSR(ioff, isrc), od, 0.85, 20)
    with gpu_ctx.from_arrays(ioff, isrc, out_degree=od) as g:
        r, _, _ = g.pagerank(iterations=20, method=_abi.PR_TILED, tile_src_bits=sb, tile_dst_bits=db)
    assert np.abs(r.astype(np.float64) - ref).sum() <= TOL_L1, (sb, db)


def test_pr_er_2p22_normalised(gpu_ctx):
    """Erdos-Renyi with V = 2^22 and 2^24 undirected pairs (2^25 directed
    edges): the size where FP32 ranks and sums start to lose digits; every
    schedule stays within the L1 bound of the FP64 oracle."""
    h = A.erdos_renyi(1 << 22, 1 << 24, seed=5)
    off, dst = h.offsets().copy(), h.dst().copy()
    V = len(off) - 1
    od = np.diff(off).astype(np.uint32)
    # symmetric: the out-CSR is also the in-CSR
    ref, _ = O.pr_norm(O.CSR(off, dst), od, 0.85, 20)
    with gpu_ctx.from_arrays(off, dst, out_degree=od) as g:
        for method in (_abi.PR_PULL, _abi.PR_TILED, _abi.PR_AUTO):
            r, _, st = g.pagerank(iterations=20, method=method)
            l1 = float(np.abs(r.astype(np.float64) - ref).sum())
            print("method", method, "L1", l1, flush=True)
            assert l1 <= TOL_L1, method
    h.free()


def test_pr_blocked_needs_resident_graph():
    off, dst, _ = G.random_graph(3000, 12000, 33, symmetric=True)
    ioff, isrc, od = csc_of(off, dst, 3000)
    E = int(ioff[-1])
    with A.Context(edge_budget_bytes=4 * (E // 3), part_bytes=4 * 4096) as ctx, \
            ctx.from_arrays(ioff, isrc, out_degree=od) as g:
        for method in (_abi.PR_BLOCKED, _abi.PR_TILED):
            with pytest.raises(A.AscError) as e:
                g.pagerank(iterations=5, method=method)
            assert e.value.code == _abi.ERR_UNSUPPORTED
        ref, _ = O.pr_norm(O.CSR(ioff, isrc), od, 0.85, 5)
        r, _, _ = g.pagerank(iterations=5)   # AUTO: the streamed pull
        assert np.abs(r.astype(np.float64) - ref).sum() <= TOL_L1


def test_pr_rmat_csc(gpu_ctx):
    h = A.rmat(14, 16, seed=4, csc=True)
    ioff, isrc, od = h.offsets().copy(), h.dst().copy(), h.out_degree().copy()
    ref, _ = O.pr_norm(O.CSR(ioff, isrc), od, 0.85, 20)
    with gpu_ctx.adopt(h) as g:
        r, _, _ = g.pagerank(iterations=20)
    assert np.abs(r.astype(np.float64) - ref).sum() <= TOL_L1


@pytest.mark.parametrize("mode", [_abi.PR_NORMALISED, _abi.PR_REFERENCE])
def test_pr_partially_resident(mode):
    """A third of the in-edges resident, the rest streamed in vertex-range
    chunks every iteration (prSumKernel_dynamic, gpu_kernels.cuh:554-575)."""
    V = 20000
    off, dst, _ = G.random_graph(V, 200000, 5, symmetric=False)
    ioff, isrc, od = csc_of(off, dst, V)
    E = len(isrc)
    with A.Context(edge_budget_bytes=4 * (E // 3), part_bytes=4 * 8192) as ctx:
        with ctx.from_arrays(ioff, isrc, out_degree=od) as g:
            info = g.info()
            assert 0 < info["static_vertices"] < V
            if mode == _abi.PR_NORMALISED:
                ref, _ = O.pr_norm(O.CSR(ioff, isrc), od, 0.85, 20)
                r, _, st = g.pagerank(iterations=20)
                assert np.abs(r.astype(np.float64) - ref).sum() <= TOL_L1
                assert st["streamed_edges"] == info["ondemand_edges"] * 20
            else:
                ref, rst = O.pr_ref(O.CSR(ioff, isrc), od)
                r, _, st = g.pagerank(mode=_abi.PR_REFERENCE)
                assert st["iterations"] == rst["iterations"]
                assert np.array_equal(r, ref)


@pytest.mark.parametrize("symmetric", [True, False])
def test_pr_reference_mode_exact(gpu_ctx, symmetric):
    V = 3000
    off, dst, _ = G.random_graph(V, 15000, 7, symmetric=symmetric)
    ioff, isrc, od = csc_of(off, dst, V)
    ref, rst = O.pr_ref(O.CSR(ioff, isrc), od)
    with gpu_ctx.from_arrays(ioff, isrc, out_degree=od) as g:
        r, _, st = g.pagerank(mode=_abi.PR_REFERENCE)
    assert st["iterations"] == rst["iterations"]
    assert np.array_equal(r, ref)
