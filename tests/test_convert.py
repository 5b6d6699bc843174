"""Edge-list conversion (asc_read_edge_list, ptgraph --type convert) on the CPU.

The reference converter (converter/converter.cpp) has no tests; these pin the
conversion against CSRs built independently with numpy (tests/graphs.py) and
check that the reference converter's recorded defects are absent: V taken
from the sources only (converter.cpp:126), trailing empty vertices, zero
weights (converter.cpp:308, 396)."""
import os
import subprocess

import numpy as np
import pytest

import ascetic_amd as A
from ascetic_amd import _abi
from tests import graphs as G


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def as_arrays(h):
    off = np.array(h.offsets(), dtype=np.uint64)
    return off, np.array(h.dst(), dtype=np.uint32), (None if h.weights() is None else np.array(h.weights()))


def test_text_with_comments_and_separators(tmp_path):
    edges = [(0, 3), (3, 1), (2, 0), (0, 3), (4, 4), (1, 2)]
    text = "# a comment\n% another\n\n" + "".join(
        f"{u}{sep}{v}\n" for (u, v), sep in zip(edges, [" ", "\t", ",", "  ", " ", "\t "]))
    p = write(tmp_path, "g.txt", text)
    off, dst, _ = as_arrays(A.read_edge_list(p))
    eo, ed, _ = G.csr_from_edges(5, [e for e in edges if e[0] != e[1]])
    assert np.array_equal(off, eo) and np.array_equal(dst, ed)
    # duplicates and self-loops kept when asked
    off2, dst2, _ = as_arrays(A.read_edge_list(p, dedup=False, keep_self_loops=True))
    eo2, ed2, _ = G.csr_from_edges(5, edges, dedup=False)
    assert np.array_equal(off2, eo2) and np.array_equal(dst2, ed2)


def test_vertices_from_both_sides_and_trailing_isolated(tmp_path):
    # the largest id appears only as a destination (the reference's V would drop it)
    p = write(tmp_path, "g.txt", "0 1\n1 9\n2 0\n")
    h = A.read_edge_list(p)
    assert h.num_vertices == 10
    h2 = A.read_edge_list(p, num_vertices=14)
    off, dst, _ = as_arrays(h2)
    assert h2.num_vertices == 14 and off[-1] == 3 and np.all(off[10:] == 3)
    with pytest.raises(A.AscError) as e:
        A.read_edge_list(p, num_vertices=5)
    assert e.value.code == _abi.ERR_RANGE


def test_symmetrize_matches_numpy(tmp_path):
    rng = np.random.default_rng(4)
    u = rng.integers(0, 300, 2000)
    v = rng.integers(0, 300, 2000)
    p = write(tmp_path, "g.txt", "".join(f"{a} {b}\n" for a, b in zip(u, v)))
    off, dst, _ = as_arrays(A.read_edge_list(p, symmetrize=True, num_vertices=300))
    eo, ed, _ = G.csr_from_edges(300, [(a, b) for a, b in zip(u.tolist(), v.tolist()) if a != b], symmetric=True)
    assert np.array_equal(off, eo) and np.array_equal(dst, ed)


def test_weights_given_and_generated(tmp_path):
    p = write(tmp_path, "w.txt", "0 1 7\n1 0 9\n1 2 3\n0 1 5\n")
    off, dst, w = as_arrays(A.read_edge_list(p, weighted=True))
    # dedup keeps the lightest copy of (0, 1)
    assert off.tolist() == [0, 1, 3, 3] and dst.tolist() == [1, 0, 2] and w.tolist() == [5, 9, 3]
    rng = np.random.default_rng(5)
    u, v = rng.integers(0, 500, 3000), rng.integers(0, 500, 3000)
    q = write(tmp_path, "g.txt", "".join(f"{a} {b}\n" for a, b in zip(u, v)))
    h1 = A.read_edge_list(q, symmetrize=True, weighted=True, wmin=1, wmax=20, seed=3)
    off, dst, w = as_arrays(h1)
    assert w.min() >= 1 and w.max() <= 20 and len(np.unique(w)) == 20
    # w(u, v) = w(v, u), and the same seed gives the same weights
    src = np.repeat(np.arange(len(off) - 1), np.diff(off).astype(np.int64))
    lookup = {(int(a), int(b)): int(c) for a, b, c in zip(src, dst, w)}
    assert all(lookup[(b, a)] == c for (a, b), c in lookup.items())
    _, _, w2 = as_arrays(A.read_edge_list(q, symmetrize=True, weighted=True, wmin=1, wmax=20, seed=3))
    assert np.array_equal(w, w2)
    with pytest.raises(A.AscError) as e:
        A.read_edge_list(write(tmp_path, "mixed.txt", "0 1 4\n1 2\n"), weighted=True)
    assert e.value.code == _abi.ERR_FORMAT


def test_matrix_market_one_based(tmp_path):
    text = "%%MatrixMarket matrix coordinate pattern general\n% comment\n4 4 3\n1 2\n2 3\n4 1\n"
    off, dst, _ = as_arrays(A.read_edge_list(write(tmp_path, "g.mtx", text)))
    eo, ed, _ = G.csr_from_edges(4, [(0, 1), (1, 2), (3, 0)])
    assert np.array_equal(off, eo) and np.array_equal(dst, ed)
    with pytest.raises(A.AscError) as e:
        A.read_edge_list(write(tmp_path, "z.txt", "0 1\n"), one_based=True)
    assert e.value.code == _abi.ERR_FORMAT


def test_binary_pairs_and_errors(tmp_path):
    pairs = np.array([[0, 2], [2, 1], [1, 0], [3, 2]], dtype=np.uint32)
    p = tmp_path / "g.toB"
    pairs.tofile(p)
    off, dst, _ = as_arrays(A.read_edge_list(str(p)))
    eo, ed, _ = G.csr_from_edges(4, pairs.tolist())
    assert np.array_equal(off, eo) and np.array_equal(dst, ed)
    (tmp_path / "bad.toB").write_bytes(b"\x00" * 12)
    with pytest.raises(A.AscError) as e:
        A.read_edge_list(str(tmp_path / "bad.toB"))
    assert e.value.code == _abi.ERR_FORMAT
    for bad in ["0 x\n", "0\n", "1 99999999999\n"]:
        with pytest.raises(A.AscError) as e:
            A.read_edge_list(write(tmp_path, "b.txt", bad))
        assert e.value.code == _abi.ERR_FORMAT, bad


def test_csc_output_is_transpose_with_out_degrees(tmp_path):
    rng = np.random.default_rng(6)
    u, v = rng.integers(0, 200, 1500), rng.integers(0, 200, 1500)
    p = write(tmp_path, "g.txt", "".join(f"{a} {b}\n" for a, b in zip(u, v)))
    h = A.read_edge_list(p, csc=True, num_vertices=200)
    ioff, isrc, _ = as_arrays(h)
    oo, od, _ = G.csr_from_edges(200, [(a, b) for a, b in zip(u.tolist(), v.tolist()) if a != b])
    to, ts = G.transpose(oo, od, 200)
    assert np.array_equal(ioff, to) and np.array_equal(isrc, ts)
    assert np.array_equal(np.array(h.out_degree()), np.diff(oo).astype(np.uint32))


def test_many_lines_parse_in_parallel_chunks(tmp_path):
    rng = np.random.default_rng(7)
    u, v = rng.integers(0, 5000, 200000), rng.integers(0, 5000, 200000)
    p = write(tmp_path, "big.txt", "".join(f"{a}\t{b}\n" for a, b in zip(u, v)))
    a = as_arrays(A.read_edge_list(p, threads=7))
    b = as_arrays(A.read_edge_list(p, threads=1))
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    eo, ed, _ = G.csr_from_edges(5000, [(x, y) for x, y in zip(u.tolist(), v.tolist()) if x != y])
    assert np.array_equal(a[0], eo) and np.array_equal(a[1], ed)


def test_cli_convert_round_trip(tmp_path):
    p = write(tmp_path, "g.txt", "0 1\n1 2\n2 0\n5 1\n")
    out = str(tmp_path / "g.bwcsr")
    r = subprocess.run([A._abi.CLI_PATH, "--type", "convert", "--input", p, "--out", out, "--format", "bwcsr",
                        "--symmetrize", "1", "--wmin", "2", "--wmax", "9"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    h = A.read_graph(out)
    off, dst, w = as_arrays(h)
    eo, ed, _ = G.csr_from_edges(6, [(0, 1), (1, 2), (2, 0), (5, 1)], symmetric=True)
    assert np.array_equal(off, eo) and np.array_equal(dst, ed)
    assert w.min() >= 2 and w.max() <= 9
