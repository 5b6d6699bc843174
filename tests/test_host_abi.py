"""Host-side C ABI checks that need no GPU: exported symbols, the static /
on-demand planner against the oracle's restatement of getMaxPartitionSize,
part splitting, file formats, the generator and the CLI."""
import os
import re
import subprocess

import numpy as np
import pytest

import ascetic_amd as A
from ascetic_amd import _abi
from oracle import oracle as O
from tests import graphs as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "ascetic.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(asc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = _abi.load()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python mirror binds exactly the declared set
    assert set(_abi.SIGNATURES) == set(names)
    assert lib.asc_abi_version() == 1
    assert lib.asc_strerror(_abi.ERR_FORMAT).decode().startswith("file does not match")


def test_no_gpu_context_fails_cleanly():
    if A.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(A.AscError) as e:
        A.Context()
    assert e.value.code == _abi.ERR_NOGPU


# ---- planning ------------------------------------------------------------------

@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("edge_size,param", [(4, 13), (8, 15)])
def test_plan_matches_reference_rule(seed, edge_size, param):
    rng = np.random.default_rng(seed)
    V = 5000
    deg = rng.integers(1, 40, V)  # no zero-degree vertex: both rules flag the same prefix
    off = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    E = int(off[-1])
    for total_edges in (E // 7, E // 3, E, 3 * E, 100 * E):
        free = param * 4 * V + edge_size * total_edges
        for rate in (0.0, 0.3, 0.8):
            ref = O.plan_reference(off, edge_size, free, param, advise_rate=rate)
            ours = A.plan_partition(off, edge_size, free, param, advise_rate=rate)
            assert ours["total_device_edges"] == ref["total_gpu_size"]
            assert abs(ours["advise_k"] - ref["advise_k"]) < 1e-6
            if ref["advise_k"] * ref["total_gpu_size"] < E:
                # same floored prefix, same static vertices and edges
                assert ours["static_vertices"] == ref["static_vertices"]
                assert ours["static_edges"] == E - ref["overload_size"]
                assert ours["ondemand_edges"] == ref["overload_size"]
            else:
                # a prefix covering every edge is not floored to 4096 edges
                # (the reference floors it and streams the tail): fully resident
                assert ours["static_edges"] == E and ours["static_vertices"] == V
                assert ref["max_partition_size"] == E - E % 4096


def test_plan_budget_and_hub_boundary():
    off = np.array([0, 10000, 10010, 10020, 20000, 20001], dtype=np.uint64)
    E = 20001
    # budget of 12000 edges -> floored to 8192: the first vertex already crosses it
    p = A.plan_partition(off, 4, 1 << 40, 13, edge_budget_bytes=4 * 12000)
    assert p["static_vertices"] == 0 and p["static_edges"] == 0 and p["ondemand_edges"] == E
    p = A.plan_partition(off, 4, 1 << 40, 13, edge_budget_bytes=4 * 16384)
    assert p["static_vertices"] == 3 and p["static_edges"] == 10020
    p = A.plan_partition(off, 4, 1 << 40, 13, edge_budget_bytes=4 * E)
    assert p["static_vertices"] == 5 and p["ondemand_edges"] == 0
    p = A.plan_partition(off, 8, 1 << 40, 15, part_bytes=8 * 5000)
    assert p["part_edges"] == 5000


def test_split_parts_edge_granular():
    rng = np.random.default_rng(4)
    deg = rng.integers(1, 100, 1000)
    deg[500] = 50000  # one vertex larger than a part
    prefix = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    starts = A.split_parts(prefix, 4096)
    total = int(prefix[-1])
    assert starts[0] == 0 and starts[-1] == total
    sizes = np.diff(starts.astype(np.int64))
    assert np.all(sizes > 0) and np.all(sizes <= 4096)
    assert len(sizes) == -(-total // 4096)
    # the reference's vertex-granular split never terminates on this input
    assert O.split_reference(prefix[:-1], deg, 4096) is None


# ---- formats -------------------------------------------------------------------

@pytest.mark.parametrize("fmt", [_abi.FMT_BCSR, _abi.FMT_LEGACY32])
def test_unweighted_format_roundtrip(tmp_path, fmt):
    off, dst, _ = G.random_graph(300, 900, 5, symmetric=False)
    p = str(tmp_path / "g.bin")
    A.write_graph(p, fmt, off[:-1], dst, _abi.KIND_CSR)
    size = os.path.getsize(p)
    V, E = 300, len(dst)
    assert size == (16 + 8 * V + 4 * E if fmt == _abi.FMT_BCSR else 8 + 4 * V + 4 * E)
    h = A.read_graph(p)  # layout detected by file size
    assert h.kind == _abi.KIND_CSR
    assert np.array_equal(h.offsets(), off) and np.array_equal(h.dst(), dst)


@pytest.mark.parametrize("fmt", [_abi.FMT_BWCSR, _abi.FMT_LEGACY32W])
def test_weighted_format_roundtrip(tmp_path, fmt):
    off, dst, w = G.random_graph(200, 800, 6, symmetric=True, weighted=True)
    p = str(tmp_path / "g.bin")
    A.write_graph(p, fmt, off[:-1], G.wedges(dst, w), _abi.KIND_WCSR)
    h = A.read_graph(p)
    assert h.kind == _abi.KIND_WCSR
    assert np.array_equal(h.offsets(), off)
    assert np.array_equal(h.dst(), dst) and np.array_equal(h.weights(), w)


def test_bcsc_roundtrip(tmp_path):
    off, dst, _ = G.random_graph(200, 800, 7, symmetric=False)
    ioff, isrc = G.transpose(off, dst, 200)
    od = np.diff(off).astype(np.uint32)
    p = str(tmp_path / "g.bcsc")
    A.write_graph(p, _abi.FMT_BCSC, ioff[:-1], isrc, _abi.KIND_CSC, out_degree=od)
    assert os.path.getsize(p) == 16 + 12 * 200 + 4 * len(isrc)
    h = A.read_graph(p)
    assert h.kind == _abi.KIND_CSC
    assert np.array_equal(h.out_degree(), od) and np.array_equal(h.dst(), isrc)


def test_auto_format_bcsc_with_v_equal_e(tmp_path):
    """A .bcsc with V == E has the size of a .bwcsr with the same header; the
    automatic layout choice keeps trying size-matching layouts until one
    validates, so the file still reads as the in-CSR it is."""
    V = 64
    off = np.arange(V + 1, dtype=np.uint64)                 # a directed cycle: one out-edge each
    dst = ((np.arange(V) + 1) % V).astype(np.uint32)
    ioff, isrc = G.transpose(off, dst, V)
    od = np.ones(V, np.uint32)
    p = str(tmp_path / "cycle.bcsc")
    A.write_graph(p, _abi.FMT_BCSC, ioff[:-1], isrc, _abi.KIND_CSC, out_degree=od)
    assert os.path.getsize(p) == 16 + 8 * V + 8 * V        # = the BWCSR size for V == E
    h = A.read_graph(p)
    assert h.kind == _abi.KIND_CSC
    assert np.array_equal(h.offsets(), ioff) and np.array_equal(h.dst(), isrc)
    assert np.array_equal(h.out_degree(), od)


def test_reader_rejects_bad_files(tmp_path):
    off, dst, _ = G.path_graph(10)
    p = str(tmp_path / "g.bcsr")
    A.write_graph(p, _abi.FMT_BCSR, off[:-1], dst, _abi.KIND_CSR)
    raw = open(p, "rb").read()
    open(p, "wb").write(raw[:-3])  # truncated
    with pytest.raises(A.AscError) as e:
        A.read_graph(p)
    assert e.value.code == _abi.ERR_FORMAT
    bad = bytearray(raw)
    bad[-4:] = (1000).to_bytes(4, "little")  # an id >= V
    open(p, "wb").write(bytes(bad))
    with pytest.raises(A.AscError) as e:
        A.read_graph(p, _abi.FMT_BCSR)
    assert e.value.code == _abi.ERR_FORMAT
    with pytest.raises(A.AscError) as e:
        A.read_graph(str(tmp_path / "missing.bcsr"))
    assert e.value.code == _abi.ERR_IO


# ---- generator -----------------------------------------------------------------

def test_rmat_generator_properties():
    h = A.rmat(14, 16, seed=9, weighted=True, wmin=1, wmax=64, threads=3)
    h2 = A.rmat(14, 16, seed=9, weighted=True, wmin=1, wmax=64, threads=8)
    off, ew = h.offsets(), h.edge_words()
    assert np.array_equal(off, h2.offsets()) and np.array_equal(ew, h2.edge_words())  # thread-count independent
    V = 1 << 14
    src = np.repeat(np.arange(V), np.diff(off).astype(np.int64))
    dst, w = ew[:, 0].astype(np.int64), ew[:, 1]
    assert np.all(src != dst)                                  # no self-loops
    key = src * V + dst
    assert np.all(np.diff(key) > 0)                            # sorted, no duplicates
    assert np.array_equal(np.sort(dst * V + src), key)         # every (u,v) has its (v,u)
    rev = dict(zip(key.tolist(), w.tolist()))
    assert all(rev[int(d) * V + int(s)] == ww for s, d, ww in zip(src[:5000], dst[:5000], w[:5000]))  # symmetric
    assert w.min() >= 1 and w.max() <= 64
    deg = np.diff(off)
    assert int(np.argmax(deg)) == 0                             # no permutation: vertex 0 is the top hub
    assert A.rmat(14, 16, seed=10).num_edges != h.num_edges or not np.array_equal(
        A.rmat(14, 16, seed=10).dst(), h.dst())


def test_er_and_csc_generator():
    h = A.erdos_renyi(5000, 20000, seed=3, symmetrize=False, csc=True)
    ioff, isrc, od = h.offsets(), h.dst(), h.out_degree()
    V = 5000
    assert od.sum() == h.num_edges == ioff[-1]
    # in-CSR is the transpose of the out-CSR implied by (src -> dst)
    dst_of_in = np.repeat(np.arange(V), np.diff(ioff).astype(np.int64))
    outdeg = np.bincount(isrc, minlength=V)
    assert np.array_equal(outdeg, od)
    assert np.all(isrc != dst_of_in)


def test_rmat_scale20_matches_reference_shape():
    """R-MAT s20 ef16 symmetrized: ~31.4M directed edges, ~0.4M isolated
    vertices, vertex 0 the hub with degree ~64.6K."""
    h = A.rmat(20, 16, seed=1)
    deg = np.diff(h.offsets())
    assert 31.0e6 < h.num_edges < 31.8e6
    assert 0.38e6 < int((deg == 0).sum()) < 0.42e6
    assert int(np.argmax(deg)) == 0 and 62000 < int(deg[0]) < 67000


# ---- CLI -------------------------------------------------------------------------

def test_cli_generate_and_no_gpu(tmp_path):
    cli = _abi.CLI_PATH
    out = str(tmp_path / "g.bwcsr")
    r = subprocess.run([cli, "--type", "gen", "--gen", "rmat", "--scale", "8", "--ef", "4", "--weighted", "1",
                        "--out", out, "--format", "bwcsr"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    h = A.read_graph(out)
    assert h.kind == _abi.KIND_WCSR and h.num_vertices == 256
    r = subprocess.run([cli, "--input", out, "--type", "sssp", "--source", "0"], capture_output=True, text=True)
    if A.device_count() == 0:
        assert r.returncode == 1 and "no HIP device" in r.stderr
    r = subprocess.run([cli], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
