"""The oracle's independent restatement of the library's R-MAT sampler
(orc_rmat_pairs), which the full-scale benchmark uses to draw edges of a
graph too large for a host copy, is pinned against the host generator: over
all samples of a small graph, the pairs in both directions (self-loops
dropped) are exactly the generated CSR's edge set and carry its weights; a
spread subset of a larger graph's samples all lie in its CSR."""
import numpy as np

import ascetic_amd as A
from oracle import oracle as O


def edge_set(off, dst):
    src = np.repeat(np.arange(len(off) - 1, dtype=np.int64), np.diff(off).astype(np.int64))
    return src * (len(off) - 1) + dst.astype(np.int64)


def test_all_samples_are_the_csr():
    scale, ef, seed = 10, 8, 11
    h = A.rmat(scale, ef, seed=seed, weighted=True, wmin=1, wmax=255)
    off, ew = h.offsets().copy(), h.edge_words().copy()
    V, M = 1 << scale, ef << scale
    u, v, w = O.rmat_pairs(seed, scale, M, M, wmin=1, wmax=255)
    keep = u != v
    u, v, w = u[keep].astype(np.int64), v[keep].astype(np.int64), w[keep]
    both = np.unique(np.concatenate([u * V + v, v * V + u]))
    got = edge_set(off, ew[:, 0])
    assert np.array_equal(both, got)
    # weights: the CSR's weight of each sampled pair, both directions
    pos = np.searchsorted(got, u * V + v)
    assert np.array_equal(ew[pos, 1], w)
    pos = np.searchsorted(got, v * V + u)
    assert np.array_equal(ew[pos, 1], w)
    assert w.min() >= 1 and w.max() <= 255


def test_spread_samples_lie_in_the_csr():
    scale, ef, seed = 16, 16, 24
    h = A.rmat(scale, ef, seed=seed, weighted=True, wmin=1, wmax=64)
    off, ew = h.offsets().copy(), h.edge_words().copy()
    V, M = 1 << scale, ef << scale
    u, v, w = O.rmat_pairs(seed, scale, M, 50000, wmin=1, wmax=64)
    keep = u != v
    u, v, w = u[keep].astype(np.int64), v[keep].astype(np.int64), w[keep]
    got = edge_set(off, ew[:, 0])
    pos = np.searchsorted(got, u * V + v)
    assert np.array_equal(got[pos], u * V + v)
    assert np.array_equal(ew[pos, 1], w)
