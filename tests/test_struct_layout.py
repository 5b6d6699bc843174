"""The ctypes mirrors in ascetic_amd/_abi.py match include/ascetic.h field by
field: a C program compiled against the header prints every struct's size and
every mirrored field's offset, which must equal what ctypes computes."""
import os
import shutil
import subprocess

import pytest

from ascetic_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PAIRS = {
    "Config": "asc_config", "CsrView": "asc_csr_view", "GraphInfo": "asc_graph_info",
    "KernelStats": "asc_kernel_stats", "Stats": "asc_stats", "PrOpts": "asc_pr_opts", "Plan": "asc_plan",
    "GenParams": "asc_gen_params", "EdgeListOpts": "asc_edge_list_opts", "HostCsr": "asc_host_csr",
    "ServerOpts": "asc_server_opts", "ServerStats": "asc_server_stats", "QStats": "asc_qstats",
    "Tuning": "asc_tuning",
}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_ctypes_mirrors_match_header(tmp_path):
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "ascetic.h"', "int main(void) {"]
    expect = []
    for py, c in PAIRS.items():
        cls = getattr(_abi, py)
        lines.append(f'    printf("%zu\\n", sizeof({c}));')
        expect.append((f"sizeof({c})", __import__("ctypes").sizeof(cls)))
        for name, _ in cls._fields_:
            lines.append(f'    printf("%zu\\n", offsetof({c}, {name}));')
            expect.append((f"{c}.{name}", getattr(cls, name).offset))
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert len(got) == len(expect)
    bad = [(what, want, have) for (what, want), have in zip(expect, got) if want != have]
    assert not bad, bad
