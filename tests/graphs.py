"""Small graphs for the tests, as numpy CSR (V+1 offsets, sorted adjacency)."""
import numpy as np


def csr_from_edges(V, edges, weights=None, symmetric=False, dedup=True):
    """edges: iterable of (u, v).  Returns (off[V+1] u64, dst u32, w u32|None).
    With symmetric=True, (v, u) is added for every (u, v) (same weight)."""
    e = np.asarray(list(edges), dtype=np.int64).reshape(-1, 2)
    w = None if weights is None else np.asarray(list(weights), dtype=np.int64)
    if symmetric and len(e):
        e = np.concatenate([e, e[:, ::-1]])
        if w is not None:
            w = np.concatenate([w, w])
    if len(e):
        if w is None:
            order = np.lexsort((e[:, 1], e[:, 0]))
            e = e[order]
            if dedup:
                keep = np.ones(len(e), bool)
                keep[1:] = np.any(e[1:] != e[:-1], axis=1)
                e = e[keep]
        else:
            order = np.lexsort((w, e[:, 1], e[:, 0]))
            e, w = e[order], w[order]
            if dedup:  # keep the lightest copy of each (u, v)
                keep = np.ones(len(e), bool)
                keep[1:] = np.any(e[1:] != e[:-1], axis=1)
                e, w = e[keep], w[keep]
    deg = np.bincount(e[:, 0], minlength=V) if len(e) else np.zeros(V, np.int64)
    off = np.zeros(V + 1, dtype=np.uint64)
    off[1:] = np.cumsum(deg)
    dst = e[:, 1].astype(np.uint32) if len(e) else np.zeros(0, np.uint32)
    return off, dst, (None if w is None else w.astype(np.uint32))


def random_graph(V, M, seed, symmetric=True, weighted=False, wmax=64, self_loops=False):
    rng = np.random.default_rng(seed)
    u = rng.integers(0, V, M)
    v = rng.integers(0, V, M)
    if not self_loops:
        keep = u != v
        u, v = u[keep], v[keep]
    w = rng.integers(1, wmax + 1, len(u)) if weighted else None
    return csr_from_edges(V, zip(u.tolist(), v.tolist()), None if w is None else w.tolist(), symmetric=symmetric)


def transpose(off, dst, V):
    """In-CSR of an out-CSR (sources sorted within each in-list)."""
    src = np.repeat(np.arange(V, dtype=np.int64), np.diff(off).astype(np.int64))
    order = np.lexsort((src, dst))
    ioff = np.zeros(V + 1, dtype=np.uint64)
    ioff[1:] = np.cumsum(np.bincount(dst, minlength=V))
    return ioff, src[order].astype(np.uint32)


def wedges(dst, w):
    """AoS {dst, w} records as an (E, 2) u32 array."""
    return np.stack([np.asarray(dst, np.uint32), np.asarray(w, np.uint32)], axis=1).copy()


# Hand graphs with known answers ------------------------------------------------

def path_graph(n):
    return csr_from_edges(n, [(i, i + 1) for i in range(n - 1)], symmetric=True)


def star_graph(n):
    return csr_from_edges(n, [(0, i) for i in range(1, n)], symmetric=True)


def two_components(n1, n2):
    e = [(i, i + 1) for i in range(n1 - 1)] + [(n1 + i, n1 + i + 1) for i in range(n2 - 1)]
    return csr_from_edges(n1 + n2, e, symmetric=True)
