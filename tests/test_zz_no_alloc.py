"""asc_bfs_noalloc through the C ABI against asc_bfs (the entry point it is a
variant of): two calls in a row on one device, the second on a smaller
graph, each bit-identical to asc_bfs; and the documented refusals.  Device
buffers for the results come from the HIP runtime the library itself links
(ctypes), so the test adds no other GPU runtime to the process."""
import ctypes as C
import sys

import numpy as np
import pytest

import ascetic_amd as A
from ascetic_amd import _abi
from oracle import oracle as O
from tests import graphs as G

pytestmark = pytest.mark.gpu


def _hip():
    _abi.load()
    # the HIP runtime libascetic.so was linked against (already mapped)
    path = "libamdhip64.so.7"
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so" in line:
                path = line.split()[-1]
                break
    lib = C.CDLL(path)
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return lib


HIP_D2H = 2


class DevBuf:
    def __init__(self, hip, n):
        self.hip, self.n, self.p = hip, n, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), 4 * n) == 0
        assert hip.hipMemset(self.p, 0xAB, 4 * n) == 0     # stale contents the call must overwrite

    def get(self):
        out = np.empty(self.n, np.uint32)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, 4 * self.n, HIP_D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _noalloc(g, src, buf):
    return _abi.load().asc_bfs_noalloc(g._h, src, buf.p, None)


def test_noalloc_matches_bfs_twice_second_smaller():
    hip = _hip()
    big = A.rmat(17, 16, seed=11)                      # 131072 vertices, a hub above the heavy threshold
    boff, bdst = big.offsets().copy(), big.dst().copy()
    small_off, small_dst, _ = G.random_graph(3000, 9000, 12, symmetric=False)
    with A.Context() as ctx:
        with ctx.adopt(big, offset_bytes=4) as gb, ctx.from_arrays(small_off, small_dst) as gs:
            bufs = []
            for g, src, V in ((gb, 0, len(boff) - 1), (gs, 5, 3000)):
                print("noalloc bfs V", V, "src", src, flush=True)
                sys.stdout.flush()
                buf = DevBuf(hip, V)
                bufs.append(buf)
                assert _noalloc(g, src, buf) == 0
                ref, _ = g.bfs(src)
                assert np.array_equal(buf.get(), ref)
            # the first result is untouched by the second call
            ref, _ = gb.bfs(0)
            assert np.array_equal(bufs[0].get(), ref)
            assert np.array_equal(ref, O.bfs_queue(O.CSR(boff, bdst), 0))
            for b in bufs:
                b.free()


def test_noalloc_weighted_and_zero_based():
    hip = _hip()
    off, dst, w = G.random_graph(20000, 80000, 14, symmetric=True, weighted=True)
    with A.Context(value_convention=_abi.VALUES_ZERO_BASED) as ctx, ctx.from_arrays(off, G.wedges(dst, w)) as g:
        buf = DevBuf(hip, 20000)
        for src in (0, 19999):
            assert _noalloc(g, src, buf) == 0
            ref, _ = g.bfs(src)
            assert np.array_equal(buf.get(), ref)
        buf.free()


def test_noalloc_refusals():
    hip = _hip()
    off, dst, _ = G.random_graph(3000, 12000, 15, symmetric=True)
    E = int(off[-1])
    buf = DevBuf(hip, 3000)
    with A.Context(edge_budget_bytes=4 * (E // 3), part_bytes=4 * 4096) as ctx, ctx.from_arrays(off, dst) as g:
        assert _noalloc(g, 0, buf) == _abi.ERR_UNSUPPORTED     # not fully resident
    with A.Context() as ctx, ctx.from_arrays(off, dst) as g:
        assert _noalloc(g, 3000, buf) == _abi.ERR_INVALID       # source out of range
        assert _abi.load().asc_bfs_noalloc(g._h, 0, None, None) == _abi.ERR_INVALID
    buf.free()
    # beyond the scratch array's maximum size
    V = (1 << 21) + 1
    off = np.zeros(V + 1, np.uint64)
    off[1:] = 1
    off[0] = 0
    dst = np.array([1], np.uint32)
    big = DevBuf(hip, 1)
    with A.Context() as ctx, ctx.from_arrays(off, dst) as g:
        assert _noalloc(g, 0, big) == _abi.ERR_RANGE
    big.free()
