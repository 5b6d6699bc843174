"""The device generator (asc_generate_device) builds the host generator's CSR
bit for bit: offsets, destinations, pair weights and out-degrees, for R-MAT
and Erdos-Renyi graphs, symmetrized or not, weighted or not, as an in-CSR,
and when the device memory limit forces the vertices into several ranges."""
import sys

import numpy as np
import pytest

import ascetic_amd as A

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = [
    ("rmat14_w255", dict(type=0, scale=14, edge_factor=16, seed=3, weighted=True, wmin=1, wmax=255)),
    ("rmat12_directed", dict(type=0, scale=12, edge_factor=8, seed=1, symmetrize=False)),
    ("rmat13_csc", dict(type=0, scale=13, edge_factor=16, seed=4, csc=True)),
    ("er_w64", dict(type=1, num_vertices=100003, num_samples=500000, seed=7, weighted=True, wmin=1, wmax=64)),
    ("er_tiny", dict(type=1, num_vertices=3, num_samples=40, seed=2)),
    ("rmat1", dict(type=0, scale=1, edge_factor=4, seed=9, weighted=True)),
]


def same(h1, h2):
    assert h1.num_vertices == h2.num_vertices and h1.num_edges == h2.num_edges and h1.kind == h2.kind
    assert np.array_equal(h1.offsets(), h2.offsets())
    assert np.array_equal(h1.edge_words(), h2.edge_words())
    d1, d2 = h1.out_degree(), h2.out_degree()
    assert (d1 is None) == (d2 is None)
    if d1 is not None:
        assert np.array_equal(d1, d2)


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_device_generator_matches_host(gpu_ctx, name, kw):
    say("device-generate", name)
    same(A.generate(device=gpu_ctx, **kw), A.generate(**kw))


def test_device_generator_in_ranges(gpu_ctx):
    """A device limit that leaves room for about a third of the raw pairs at
    a time cuts the vertices into several ranges; the CSR is the same."""
    kw = dict(type=0, scale=15, edge_factor=16, seed=5, weighted=True, wmin=1, wmax=64)
    ref = A.generate(**kw)
    V = ref.num_vertices
    raw = 2 * 16 * V
    limit = 20 * V + (64 << 20) + 16 * raw // 3 * 4 // 3
    say("device-generate in ranges, limit", limit)
    with A.Context(device_mem_limit=limit, verbose=1) as ctx:
        same(A.generate(device=ctx, **kw), ref)


def test_device_generator_rejects_directed_csc(gpu_ctx):
    with pytest.raises(A.AscError):
        A.generate(device=gpu_ctx, type=0, scale=8, symmetrize=False, csc=True)
