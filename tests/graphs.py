"""Small graph builders with known triangle counts, and brute-force counters.

All builders return ``(n, src, dst)`` raw edge samples (uint32 numpy arrays);
cleaning (loops, repeats, symmetrisation) is left to the code under test.
"""
from __future__ import annotations

import itertools
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _edges(n, pairs):
    a = np.asarray(pairs, dtype=np.uint32).reshape(-1, 2)
    return n, a[:, 0].copy(), a[:, 1].copy()


def complete(n):
    """K_n: n(n-1)(n-2)/6 triangles."""
    iu, ju = np.triu_indices(n, 1)
    return n, iu.astype(np.uint32), ju.astype(np.uint32)


def wheel(rim):
    """Hub 0 plus a rim cycle of ``rim`` >= 4 vertices: rim triangles."""
    pairs = [(0, i) for i in range(1, rim + 1)]
    pairs += [(i, i % rim + 1) for i in range(1, rim + 1)]
    return _edges(rim + 1, pairs)


def windmill(k):
    """Wd(3, k): k triangles sharing vertex 0."""
    pairs = []
    for i in range(k):
        a, b = 2 * i + 1, 2 * i + 2
        pairs += [(0, a), (0, b), (a, b)]
    return _edges(2 * k + 1, pairs)


def star(k):
    return _edges(k + 1, [(0, i) for i in range(1, k + 1)])


def complete_bipartite(a, b):
    return _edges(a + b, [(i, a + j) for i in range(a) for j in range(b)])


def random_tree(n, seed=0):
    rng = np.random.default_rng(seed)
    return _edges(n, [(i, int(rng.integers(0, i))) for i in range(1, n)])


def random_multigraph(n, m, seed=0, loops=True):
    """Uniform pairs with repeats, reversed repeats and (optionally) loops."""
    rng = np.random.default_rng(seed)
    s = rng.integers(0, n, m).astype(np.uint32)
    d = rng.integers(0, n, m).astype(np.uint32)
    if not loops:
        keep = s != d
        s, d = s[keep], d[keep]
    # add reversed duplicates of a few pairs
    k = min(len(s), max(1, m // 10))
    s = np.concatenate([s, d[:k]])
    d = np.concatenate([d, s[:k]])
    return n, s, d


def toy_test_cu():
    """test_code/test.cu:47-49 CSR (row_ptr {0,2,4,6,7}, col {1,2,0,2,0,1,3})."""
    with open(os.path.join(GOLDEN, "toy_csrs.json")) as f:
        t = json.load(f)["test_cu"]
    return _csr_to_edges(t["row_ptr"], t["col"])


def toy_edge_filter_cu():
    """test_code/edge_filter.cu:31-32 CSR (first row_ptr[10] = 36 column entries)."""
    with open(os.path.join(GOLDEN, "toy_csrs.json")) as f:
        t = json.load(f)["edge_filter_cu"]
    return _csr_to_edges(t["row_ptr"], t["col"][: t["row_ptr"][-1]])


def _csr_to_edges(row_ptr, col):
    n = len(row_ptr) - 1
    pairs = [(r, col[j]) for r in range(n) for j in range(row_ptr[r], row_ptr[r + 1])]
    n = max(n, max(c for _, c in pairs) + 1)
    return _edges(n, pairs)


# ------------------------------------------------------------------ brute force
def simple_edge_set(src, dst):
    return {(int(min(a, b)), int(max(a, b))) for a, b in zip(src, dst) if a != b}


def brute_triangles(src, dst):
    es = simple_edge_set(src, dst)
    adj = {}
    for a, b in es:
        adj.setdefault(a, set()).add(b)
        adj.setdefault(b, set()).add(a)
    t = 0
    for a, b in es:
        t += len(adj[a] & adj[b])
    return t // 3


def brute_triangle_list(src, dst):
    es = simple_edge_set(src, dst)
    vs = sorted({v for e in es for v in e})
    out = []
    for a, b, c in itertools.combinations(vs, 3):
        if (a, b) in es and (a, c) in es and (b, c) in es:
            out.append((a, b, c))
    return out


def owned_in_lists(off, col, n_owned):
    """In-lists by definition: for w, owned sources u < n_owned of arcs u->w with d+(u) <= d+(w), ascending."""
    n = len(off) - 1
    d = np.diff(off.astype(np.int64))
    src = np.repeat(np.arange(n, dtype=np.int64), d)
    dst = col.astype(np.int64)
    keep = (src < n_owned) & (d[src] <= d[dst])
    s, t = src[keep], dst[keep]
    order = np.lexsort((s, t))
    s, t = s[order], t[order]
    in_off = np.zeros(n + 1, np.uint64)
    np.add.at(in_off, t + 1, 1)
    return np.cumsum(in_off).astype(np.uint64), s.astype(np.uint32)
