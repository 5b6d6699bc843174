"""The reference's own entry path, end to end on the GPU.

main.cu:11-19 -> bfs_opt / sssp_opt / cc_opt / pr_opt (CalculateOpt.cuh:11-14)
-> readDataFromFile (GraphMeta.cuh:151-171).  Graph files in every binary
layout are written with asc_write_graph and read back by asc_graph_load, then
run four ways, each compared with the oracle on the CPU:
  * the library's Python drivers with the reference's names;
  * the ptgraph CLI (the reference's flags, plus --out for the values);
  * a .bcsc / .bwcsr made by `ptgraph --type convert` from a text edge list;
  * the C++ shim printed in INTEGRATION.md, compiled here with g++ against
    include/ascetic.h and linked with libascetic.so beside a main() that
    dispatches on --type as main.cu does.
"""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import ascetic_amd as A
from ascetic_amd import _abi
from oracle import oracle as O
from tests import graphs as G

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOL_L1 = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class Files:
    """One weighted symmetric graph and one directed graph (for PageRank),
    written in every layout."""

    def __init__(self, d):
        V = 3000
        off, dst, _ = G.random_graph(V, 12000, 21, symmetric=True)
        # symmetric weights in [1, 64]: w(u, v) = w(v, u)
        src = np.repeat(np.arange(V, dtype=np.int64), np.diff(off).astype(np.int64))
        lo, hi = np.minimum(src, dst), np.maximum(src, dst)
        w = ((lo * 1000003 + hi * 7919) % 64 + 1).astype(np.uint32)
        self.V, self.off, self.dst, self.w = V, off, dst, w
        self.csr = O.CSR(off, dst, w)
        doff, ddst, _ = G.random_graph(V, 15000, 22, symmetric=False)
        self.ioff, self.isrc = G.transpose(doff, ddst, V)
        self.od = np.diff(doff).astype(np.uint32)
        self.src = int(np.argmax(np.diff(off)))
        self.path = {}
        for name, fmt, o, e, kind, od in (
                ("bcsr", _abi.FMT_BCSR, off, dst, _abi.KIND_CSR, None),
                ("legacy32", _abi.FMT_LEGACY32, off.astype(np.uint32), dst, _abi.KIND_CSR, None),
                ("bwcsr", _abi.FMT_BWCSR, off, G.wedges(dst, w), _abi.KIND_WCSR, None),
                ("legacy32w", _abi.FMT_LEGACY32W, off.astype(np.uint32), G.wedges(dst, w), _abi.KIND_WCSR, None),
                ("bcsc", _abi.FMT_BCSC, self.ioff, self.isrc, _abi.KIND_CSC, self.od)):
            p = os.path.join(d, f"g.{name}")
            A.write_graph(p, fmt, o[:-1], e, kind, out_degree=od)
            self.path[name] = p
        self.bfs = O.bfs_queue(self.csr, self.src)
        self.sssp = O.dijkstra(self.csr, self.src)
        self.cc = O.cc_minreach(self.csr)
        self.pr_ref, self.pr_ref_st = O.pr_ref(O.CSR(self.ioff, self.isrc), self.od)
        self.pr_norm, _ = O.pr_norm(O.CSR(self.ioff, self.isrc), self.od, 0.85, 20)


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    return Files(str(tmp_path_factory.mktemp("boundary")))


FORMATS = {"bcsr": _abi.FMT_BCSR, "legacy32": _abi.FMT_LEGACY32, "bwcsr": _abi.FMT_BWCSR,
           "legacy32w": _abi.FMT_LEGACY32W, "bcsc": _abi.FMT_BCSC}


@pytest.mark.parametrize("budget", [0, 1])
@pytest.mark.parametrize("explicit", [False, True])
@pytest.mark.parametrize("name", list(FORMATS))
def test_load_every_layout(files, name, explicit, budget):
    """asc_graph_load on each layout (detected by size, or named), fully
    resident and with a third of the edges resident (the rest streamed)."""
    say("load", name, "explicit", explicit, "budget", budget)
    fmt = FORMATS[name] if explicit else _abi.FMT_AUTO
    weighted = name in ("bwcsr", "legacy32w")
    E = len(files.isrc) if name == "bcsc" else len(files.dst)
    kw = {"edge_budget_bytes": (8 if weighted else 4) * (E // 3), "part_bytes": 8 * 2048} if budget else {}
    with A.Context(**kw) as ctx, ctx.load(files.path[name], fmt) as g:
        info = g.info()
        assert info["num_vertices"] == files.V and info["num_edges"] == E
        assert info["offset_bytes"] == (4 if name.startswith("legacy") else 8)
        if budget:
            assert info["ondemand_edges"] > 0
        if name == "bcsc":
            r, _, _ = g.pagerank(iterations=20)
            assert np.abs(r.astype(np.float64) - files.pr_norm).sum() <= TOL_L1
            r, _, st = g.pagerank(mode=_abi.PR_REFERENCE)
            assert st["iterations"] == files.pr_ref_st["iterations"]
            assert np.array_equal(r, files.pr_ref)
            return
        lv, _ = g.bfs(files.src)
        assert np.array_equal(lv, files.bfs)
        lab, _ = g.cc()
        assert np.array_equal(lab, files.cc)
        if weighted:
            d, _ = g.sssp(files.src)
            assert np.array_equal(d, files.sssp)
        else:
            with pytest.raises(A.AscError) as e:
                g.sssp(files.src)
            assert e.value.code == _abi.ERR_UNSUPPORTED


def test_python_drivers(files):
    """bfs_opt / sssp_opt / cc_opt / pr_opt(path, ...) with the reference's
    argument meaning (adviseRate <= 0: automatic rule)."""
    lv, st = A.bfs_opt(files.path["bcsr"], files.src, 0.0, verbose=False)
    assert np.array_equal(lv, files.bfs) and st["reached"] == int((files.bfs != files.V + 1).sum())
    d, _ = A.sssp_opt(files.path["bwcsr"], files.src, 0.0, verbose=False)
    assert np.array_equal(d, files.sssp)
    lab, _ = A.cc_opt(files.path["bcsr"], 0.0, verbose=False)
    assert np.array_equal(lab, files.cc)
    r, _, st = A.pr_opt(files.path["bcsc"], 0.0, verbose=False)
    assert np.array_equal(r, files.pr_ref) and st["iterations"] == files.pr_ref_st["iterations"]
    # an advise rate below the automatic one streams part of the graph
    d, st = A.sssp_opt(files.path["legacy32w"], files.src, 1e-7, verbose=False)
    assert np.array_equal(d, files.sssp)


def _cli(*args):
    r = subprocess.run([_abi.CLI_PATH, *map(str, args)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (args, r.stdout, r.stderr)
    return r.stdout


@pytest.mark.parametrize("extra", [[], ["--budget", 20000, "--part-mb", 1], ["--adviseK", 1e-7, "--stream", "hybrid"]])
def test_cli_every_algorithm(files, tmp_path, extra):
    """ptgraph --input f --type X --source N (ArgumentParser.cu:15-48) with the
    values written by --out."""
    say("cli", extra)
    out = str(tmp_path / "vals")
    runs = (("bfs", "bcsr", files.bfs), ("bfs", "legacy32", files.bfs), ("sssp", "bwcsr", files.sssp),
            ("sssp", "legacy32w", files.sssp), ("cc", "bcsr", files.cc), ("cc", "bwcsr", files.cc))
    for algo, name, ref in runs:
        txt = _cli("--input", files.path[name], "--type", algo, "--source", files.src, "--out", out, *extra)
        assert f'"algo": "{algo}"' in txt
        got = np.fromfile(out, dtype=np.uint32)
        assert np.array_equal(got, ref), (algo, name)
    # PageRank: the in-CSR file, reference mode (FP64) by default
    _cli("--input", files.path["bcsc"], "--type", "pr", "--out", out, *extra)
    assert np.array_equal(np.fromfile(out, dtype=np.float64), files.pr_ref)
    _cli("--input", files.path["bcsc"], "--type", "pr", "--pr-mode", "norm", "--out", out, *extra)
    assert np.abs(np.fromfile(out, dtype=np.float32).astype(np.float64) - files.pr_norm).sum() <= TOL_L1
    # zero-based values
    _cli("--input", files.path["bcsr"], "--type", "bfs", "--source", files.src, "--values", "zero", "--out", out)
    zb = np.where(files.bfs == files.V + 1, 0xFFFFFFFF, files.bfs - 1).astype(np.uint32)
    assert np.array_equal(np.fromfile(out, dtype=np.uint32), zb)
    # --type defaults to bfs and --source to 0 (main.cu:7-9, ArgumentParser.cu:8)
    _cli("--input", files.path["bcsr"], "--out", out)
    assert np.array_equal(np.fromfile(out, dtype=np.uint32), O.bfs_queue(files.csr, 0))


def test_converted_files_run(tmp_path):
    """A text edge list converted by `ptgraph --type convert` (the reference's
    converter/converter.cpp step) into .bcsc and .bwcsr, then run."""
    V, rng = 2500, np.random.default_rng(31)
    u = rng.integers(0, V, 9000)
    v = rng.integers(0, V, 9000)
    keep = u != v
    u, v = u[keep], v[keep]
    u[0], v[0] = V - 1, 0        # the largest id appears, so V comes out right
    wt = rng.integers(1, 30, len(u))
    txt = str(tmp_path / "edges.txt")
    with open(txt, "w") as f:
        f.write("# src dst weight\n")
        for a, b, c in zip(u.tolist(), v.tolist(), wt.tolist()):
            f.write(f"{a}\t{b}\t{c}\n")
    bcsc, bwcsr, out = str(tmp_path / "g.bcsc"), str(tmp_path / "g.bwcsr"), str(tmp_path / "vals")
    _cli("--type", "convert", "--input", txt, "--format", "bcsc", "--out", bcsc)
    _cli("--type", "convert", "--input", txt, "--format", "bwcsr", "--out", bwcsr)
    # expected graphs built here from the same list (duplicates keep the lightest weight)
    off, dst, w = G.csr_from_edges(V, zip(u.tolist(), v.tolist()), wt.tolist())
    ioff, isrc = G.transpose(off, dst, V)
    od = np.diff(off).astype(np.uint32)
    _cli("--input", bcsc, "--type", "pr", "--out", out)
    ref, _ = O.pr_ref(O.CSR(ioff, isrc), od)
    assert np.array_equal(np.fromfile(out, dtype=np.float64), ref)
    csr = O.CSR(off, dst, w)
    s = int(np.argmax(np.diff(off)))
    _cli("--input", bwcsr, "--type", "sssp", "--source", s, "--out", out)
    assert np.array_equal(np.fromfile(out, dtype=np.uint32), O.dijkstra(csr, s))
    _cli("--input", bwcsr, "--type", "cc", "--out", out)
    assert np.array_equal(np.fromfile(out, dtype=np.uint32), O.cc_minreach(csr))


MAIN_CPP = r"""
// main() with the reference's dispatch (main.cu:4-22, ArgumentParser.cu:15-48)
#include <cstdlib>
#include <string>
void bfs_opt(std::string path, unsigned sourceNode, double adviseRate);
void cc_opt(std::string path, double adviseRate);
void sssp_opt(std::string path, unsigned sourceNode, double adviseRate);
void pr_opt(std::string path, double adviseRate);
int main(int argc, char** argv) {
    std::string input, type = "bfs";
    unsigned source = 0;
    double adviseK = 0;
    for (int i = 1; i + 1 < argc; i += 2) {
        std::string k = argv[i], v = argv[i + 1];
        if (k == "--input") input = v;
        else if (k == "--type") type = v;
        else if (k == "--source") source = (unsigned)std::strtoul(v.c_str(), nullptr, 10);
        else if (k == "--adviseK") adviseK = std::atof(v.c_str());
    }
    if (type == "bfs") bfs_opt(input, source, adviseK);
    else if (type == "cc") cc_opt(input, adviseK);
    else if (type == "sssp") sssp_opt(input, source, adviseK);
    else if (type == "pr") pr_opt(input, adviseK);
    return 0;
}
"""


def test_integration_shim_compiles_and_runs(files, tmp_path):
    """The C++ shim printed in INTEGRATION.md, unchanged, linked with the
    library: its printed checksums equal the oracle's."""
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    shim = re.search(r"```cpp\n(.*?)```", doc, re.S).group(1)
    assert "void bfs_opt(string path" in shim
    src = tmp_path / "CalculateOpt_mi355x.cpp"
    src.write_text(shim)
    main = tmp_path / "main.cpp"
    main.write_text(MAIN_CPP)
    exe = str(tmp_path / "ptGraph")
    lib_dir = os.path.dirname(_abi.LIB_PATH)
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(main), str(src),
                        "-L", lib_dir, "-lascetic", f"-Wl,-rpath,{lib_dir}", "-o", exe],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr

    def run(*args):
        p = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=120)
        assert p.returncode == 0, p.stderr
        m = re.search(r"result checksum : (\S+)", p.stdout)
        assert m and "totalProcess time is" in p.stdout, p.stdout + p.stderr
        return m.group(1), p.stdout

    ck, txt = run("--input", files.path["bcsr"], "--type", "bfs", "--source", files.src)
    assert int(ck) == int(files.bfs.astype(np.uint64).sum())
    assert f"nodeSum : {int((files.bfs != files.V + 1).sum())}" in txt   # BFS nodeSum = vertices reached
    ck, _ = run("--input", files.path["bwcsr"], "--type", "sssp", "--source", files.src)
    assert int(ck) == int(files.sssp.astype(np.uint64).sum())
    ck, _ = run("--input", files.path["bcsr"], "--type", "cc")
    assert int(ck) == int(files.cc.astype(np.uint64).sum())
    ck, _ = run("--input", files.path["bcsc"], "--type", "pr")
    seq = 0.0
    for x in files.pr_ref.tolist():   # the shim sums in index order
        seq += x
    assert float(ck) == seq
    # adviseK below the automatic rate: part of the graph streams
    ck, txt = run("--input", files.path["bwcsr"], "--type", "sssp", "--source", files.src, "--adviseK", 1e-7)
    assert int(ck) == int(files.sssp.astype(np.uint64).sum())
    assert "move overload size : 0\n" not in txt
