"""CPU tests that pin the oracle (oracle/lstc_oracle.c) before it is trusted.

Each oracle function is checked against something it cannot share a mistake
with: closed forms, brute force, networkx / scipy, the reference's own compiled
Undir2Dir tool (through committed golden vectors), and invariants.
"""
import json
import os

import networkx as nx
import numpy as np
import pytest
import scipy.sparse as sp

import graphs as G

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def count_all(O, n, s, d, compact=True):
    sym, g = O.clean_orient(n, s, d, compact)
    return sym, g, O.count(g)


# ------------------------------------------------------------------ closed forms
@pytest.mark.parametrize("k", [3, 4, 5, 10, 40, 130, 300])
def test_complete_graph(oracle, k):
    _, g, t = count_all(oracle, *G.complete(k))
    assert t == k * (k - 1) * (k - 2) // 6
    assert oracle.count_k1(g)[0] == t


@pytest.mark.parametrize("rim", [4, 5, 17, 200])
def test_wheel(oracle, rim):
    assert count_all(oracle, *G.wheel(rim))[2] == rim


@pytest.mark.parametrize("k", [1, 2, 9, 500])
def test_windmill(oracle, k):
    assert count_all(oracle, *G.windmill(k))[2] == k


@pytest.mark.parametrize("builder", [lambda: G.star(50), lambda: G.random_tree(300, 1),
                                     lambda: G.complete_bipartite(20, 30)])
def test_triangle_free(oracle, builder):
    assert count_all(oracle, *builder())[2] == 0


@pytest.mark.parametrize("rows,cols,p,seed", [(3, 3, 1.0, 1), (10, 17, 0.3, 2), (60, 50, 0.3, 9),
                                              (40, 40, 0.0, 3)])
def test_grid_two_per_diagonal(oracle, rows, cols, p, seed):
    n, s, d = oracle.gen_grid(rows, cols, p, seed)
    D = oracle.grid_num_diagonals(rows, cols, p, seed)
    _, g, t = count_all(oracle, n, s, d)
    assert t == 2 * D
    # structure: grid edges plus diagonals
    assert g.m == rows * (cols - 1) + cols * (rows - 1) + D


def test_toy_csrs(oracle):
    toys = json.load(open(os.path.join(GOLD, "toy_csrs.json")))
    _, _, t = count_all(oracle, *G.toy_test_cu())
    assert t == toys["test_cu"]["triangles_after_cleaning"]
    _, g, t = count_all(oracle, *G.toy_edge_filter_cu())
    assert t == toys["edge_filter_cu"]["triangles_after_cleaning"]
    assert g.m == 27


# ------------------------------------------------------------------ independent counters
@pytest.mark.parametrize("seed", range(6))
def test_random_vs_networkx_and_bruteforce(oracle, seed):
    n, s, d = G.random_multigraph(60 + 10 * seed, 400 + 50 * seed, seed=seed)
    _, g, t = count_all(oracle, n, s, d)
    nxg = nx.Graph()
    nxg.add_edges_from((int(a), int(b)) for a, b in zip(s, d) if a != b)
    assert t == sum(nx.triangles(nxg).values()) // 3
    assert t == len(G.brute_triangle_list(s, d))


@pytest.mark.parametrize("scale,seed", [(9, 1), (11, 2), (12, 3)])
def test_rmat_vs_scipy_trace(oracle, scale, seed):
    n, s, d = oracle.gen_rmat(scale, 16, seed=seed)
    sym, g, t = count_all(oracle, n, s, d)
    A = sp.csr_matrix((np.ones(len(sym.col), np.int64), sym.col.astype(np.int64),
                       sym.off.astype(np.int64)), shape=(sym.n, sym.n))
    assert t == int((A @ A).multiply(A).sum()) // 6


def test_k1_restatement_per_row_matches_merge(oracle):
    # two different algorithms (tiled bisection vs two-pointer merge), per row
    for seed, (n, s, d) in enumerate([oracle.gen_rmat(12, 16, seed=5),
                                      G.complete(300), oracle.gen_er(3000, 40, seed=2)]):
        _, g = oracle.clean_orient(n, s, d)
        t1, rows1 = oracle.count(g, per_row=True)
        t2, rows2 = oracle.count_k1(g)
        assert t1 == t2
        assert np.array_equal(rows1, rows2)


def test_count_invariant_under_relabel(oracle):
    n, s, d = oracle.gen_rmat(11, 16, seed=4)
    t0 = count_all(oracle, n, s, d)[2]
    perm = np.random.default_rng(0).permutation(n).astype(np.uint32)
    assert count_all(oracle, n, perm[s], perm[d])[2] == t0
    # and under a different (still acyclic) orientation: reversed id tie-break is
    # a relabel v -> n-1-v
    assert count_all(oracle, n, (n - 1 - s).astype(np.uint32), (n - 1 - d).astype(np.uint32))[2] == t0


def test_owned_rows_sum(oracle):
    _, g = oracle.clean_orient(*oracle.gen_rmat(11, 16, seed=2))
    t, rows = oracle.count(g, per_row=True)
    for cut in [0, 1, g.n // 3, g.n]:
        assert oracle.count(g, n_owned=cut) == int(rows[:cut].sum())


# ------------------------------------------------------------------ clean semantics
@pytest.mark.parametrize("compact", [True, False])
def test_clean_matches_set_semantics(oracle, compact):
    n, s, d = G.random_multigraph(500, 3000, seed=11)
    sym = oracle.clean(n, s, d, compact)
    es = G.simple_edge_set(s, d)
    verts = sorted({v for e in es for v in e}) if compact else list(range(n))
    remap = {v: i for i, v in enumerate(verts)}
    assert sym.n == len(verts)
    assert sym.m == 2 * len(es)
    want = sorted([(remap[a], remap[b]) for a, b in es] + [(remap[b], remap[a]) for a, b in es])
    rows = np.repeat(np.arange(sym.n), np.diff(sym.off).astype(np.int64))
    got = list(zip(rows.tolist(), sym.col.tolist()))
    assert got == want


def test_clean_degenerate(oracle):
    # empty input, only loops, single edge repeated both ways
    sym = oracle.clean(5, np.array([], np.uint32), np.array([], np.uint32))
    assert sym.n == 0 and sym.m == 0
    sym = oracle.clean(5, np.array([1, 2, 2], np.uint32), np.array([1, 2, 2], np.uint32), compact=False)
    assert sym.n == 5 and sym.m == 0
    sym = oracle.clean(9, np.array([7, 3, 7, 3], np.uint32), np.array([3, 7, 3, 7], np.uint32))
    assert sym.n == 2 and sym.m == 2 and sym.col.tolist() == [1, 0]  # m counts both arcs


# ------------------------------------------------------------------ orientation
def test_orientation_matches_reference_undir2dir_golden(oracle):
    """Golden vectors written by the reference's compiled preprocess_CPU/Undir2Dir.c."""
    gold = json.load(open(os.path.join(GOLD, "orient_undir2dir.json")))["cases"]
    for name, c in gold.items():
        sym = oracle.CSR(c["n"], np.array(c["sym_off"], np.uint64), np.array(c["sym_col"], np.uint32))
        g = oracle.orient(sym)
        assert g.off.tolist() == c["dir_off"], name
        assert g.col.tolist() == c["dir_col"], name
        assert oracle.count(g) == c["triangles"], name


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref",
                                                    "undir2dir")), reason="oracle/_ref not built")
def test_orientation_matches_reference_tool_live(oracle, tmp_path):
    for seed in range(3):
        sym, g = oracle.clean_orient(*oracle.gen_rmat(9, 8, seed=seed))
        ref = oracle.run_ref_undir2dir(sym, str(tmp_path))
        assert np.array_equal(ref.off, g.off)
        assert np.array_equal(ref.col, g.col)


def test_orientation_invariants(oracle):
    sym, g = oracle.clean_orient(*oracle.gen_rmat(12, 16, seed=8))
    deg = np.diff(sym.off).astype(np.int64)
    assert g.m == sym.m // 2
    rows = np.repeat(np.arange(g.n), np.diff(g.off).astype(np.int64))
    col = g.col.astype(np.int64)
    # rank (deg, -id) strictly increases along every arc => acyclic, each edge once
    assert np.all((deg[rows] < deg[col]) | ((deg[rows] == deg[col]) & (col < rows)))
    key = np.minimum(rows, col) * g.n + np.maximum(rows, col)
    assert len(np.unique(key)) == g.m
    # rows ascending
    for u in range(0, g.n, max(1, g.n // 200)):
        r = g.col[g.off[u]:g.off[u + 1]]
        assert np.all(np.diff(r.astype(np.int64)) > 0)
    # out-degree bounded by sqrt(2m) in a degree orientation
    assert np.diff(g.off).max() <= int(np.sqrt(2 * g.m)) + 1


# ------------------------------------------------------------------ halo
def brute_halo(sym, part, p, mode):
    n = sym.n
    deg = np.diff(sym.off).astype(np.int64)

    def nbrs(v):
        return sym.col[sym.off[v]:sym.off[v + 1]].tolist()

    def arc(s, d):
        return deg[s] < deg[d] or (deg[s] == deg[d] and d < s)

    P = [v for v in range(n) if part[v] == p]
    H = sorted({d for v in P for d in nbrs(v) if part[d] != p and (mode == 0 or arc(v, d))})
    verts = P + H
    loc = {v: i for i, v in enumerate(verts)}
    off, col = [0], []
    for v in verts:
        row = sorted(loc[d] for d in nbrs(v) if d in loc and arc(v, d))
        col += row
        off.append(len(col))
    return verts, len(P), off, col


@pytest.mark.parametrize("mode", [0, 1])
def test_halo_matches_definition(oracle, mode):
    sym, g = oracle.clean_orient(*G.random_multigraph(150, 700, seed=21))
    part = oracle.partition_random(sym.n, 3, seed=4)
    for p in range(3):
        h = oracle.halo_extract(sym, part, p, mode)
        verts, np_, off, col = brute_halo(sym, part, p, mode)
        assert h.n_owned == np_
        assert h.l2g.tolist() == verts
        assert h.csr.off.tolist() == off
        assert h.csr.col.tolist() == col


@pytest.mark.parametrize("k,mode", [(1, 0), (2, 0), (8, 0), (8, 1), (13, 1)])
def test_halo_parts_sum_to_total(oracle, k, mode):
    sym, g = oracle.clean_orient(*oracle.gen_rmat(12, 16, seed=3))
    total = oracle.count(g)
    part = oracle.partition_random(sym.n, k, seed=7)
    s = 0
    for p in range(k):
        h = oracle.halo_extract(sym, part, p, mode)
        s += oracle.count(h.csr, n_owned=h.n_owned)
    assert s == total


def test_partition_random(oracle):
    part = oracle.partition_random(100000, 8, seed=1)
    assert part.max() == 7 and part.min() == 0
    cnt = np.bincount(part, minlength=8)
    assert np.all(np.abs(cnt - 12500) < 600)
    assert np.array_equal(part, oracle.partition_random(100000, 8, seed=1))
    assert not np.array_equal(part, oracle.partition_random(100000, 8, seed=2))


# ------------------------------------------------------------------ generators
def test_generators_deterministic_and_in_range(oracle):
    for f in [lambda s: oracle.gen_rmat(10, 8, seed=s), lambda s: oracle.gen_er(999, 10, seed=s),
              lambda s: oracle.gen_chung_lu(2000, seed=s), lambda s: oracle.gen_grid(7, 9, 0.5, seed=s)]:
        n, a, b = f(1)
        n2, a2, b2 = f(1)
        assert np.array_equal(a, a2) and np.array_equal(b, b2)
        assert a.max() < n and b.max() < n
        _, a3, b3 = f(2)
        assert not (np.array_equal(a, a3) and np.array_equal(b, b3))


def test_rmat_quadrant_frequencies(oracle):
    # top-level bit frequencies follow a/b/c/d: P(src msb=1) = c+d, P(dst msb=1) = b+d
    n, s, d = oracle.gen_rmat(16, 4, seed=3)
    top = 1 << 15
    ps = float(np.mean(s >= top))
    pd = float(np.mean(d >= top))
    both = float(np.mean((s >= top) & (d >= top)))
    assert abs(ps - 0.24) < 0.005 and abs(pd - 0.24) < 0.005 and abs(both - 0.05) < 0.004


def test_chung_lu_shape(oracle):
    n, s, d = oracle.gen_chung_lu(200000, gamma=2.1, mean_deg=50, w_max=5e4, seed=1)
    deg = np.bincount(np.concatenate([s, d]), minlength=n)
    assert abs(deg.mean() - 50) < 3
    assert 0.5 * 5e4 < deg.max() < 1.5 * 5e4


@pytest.mark.parametrize("cut_frac", [0.0, 0.4, 1.0])
def test_owned_in_lists_helper_by_brute_force(oracle, cut_frac):
    """The in-list definition the GPU tests compare with, checked by plain loops."""
    _, g = oracle.clean_orient(*G.random_multigraph(300, 2500, seed=4))
    cut = int(g.n * cut_frac)
    off, col = g.off.astype(np.int64), g.col.astype(np.int64)
    dplus = [int(off[u + 1] - off[u]) for u in range(g.n)]
    lists = [[] for _ in range(g.n)]
    for u in range(cut):
        for i in range(off[u], off[u + 1]):
            w = int(col[i])
            if dplus[u] <= dplus[w]:
                lists[w].append(u)
    want_off = np.cumsum([0] + [len(x) for x in lists]).astype(np.uint64)
    want_col = np.array([u for x in lists for u in sorted(x)], np.uint32)
    got_off, got_col = G.owned_in_lists(g.off, g.col, cut)
    assert np.array_equal(got_off, want_off)
    assert np.array_equal(got_col, want_col)


def test_rmat_golden_counts_reproduce(oracle):
    """The committed full-size counts come from this oracle: the small scales regenerate exactly."""
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "rmat_counts.json")))
    for s in (16, 17, 18):
        n, src, dst = oracle.gen_rmat(s, 16, seed=1)
        _, g = oracle.clean_orient(n, src, dst)
        assert (g.n, g.m, oracle.count(g)) == (gold[str(s)]["n"], gold[str(s)]["m"], gold[str(s)]["triangles"])


def test_count_merge_list_is_row_subset_of_brute_force(oracle):
    """lo_count_merge_list (the CPU baseline's sampled counter): the count over a row list equals
    sum over those rows u of sum_{w in N+(u)} |N+(u) n N+(w)| by Python sets, for random subsets."""
    import numpy as np
    n, s, d = oracle.gen_rmat(10, 16, seed=12)
    _, g = oracle.clean_orient(n, s, d)
    nbr = [set(g.col[g.off[u]:g.off[u + 1]].tolist()) for u in range(g.n)]
    per = [sum(len(nbr[u] & nbr[w]) for w in nbr[u]) for u in range(g.n)]
    rng = np.random.default_rng(3)
    for size in (1, 7, 100, g.n):
        rows = np.sort(rng.choice(g.n, size=size, replace=False)).astype(np.uint32)
        assert oracle.count_rows(g.off, g.col, rows, threads=2) == sum(per[int(r)] for r in rows)
    assert oracle.count_rows(g.off, g.col, np.zeros(0, np.uint32)) == 0
    # every row once is the whole count
    assert oracle.count_rows(g.off, g.col, np.arange(g.n, dtype=np.uint32)) == oracle.count(g)


def test_stats_match_numpy(oracle):
    """lo_stats (max d+, sum d+^2, sum d- d+, rows with d+ >= 2) against numpy on the same CSR."""
    import numpy as np
    for n, s, d in (oracle.gen_rmat(11, 16, seed=2), oracle.gen_er(3000, 10, seed=1)):
        _, g = oracle.clean_orient(n, s, d)
        dout = np.diff(g.off.astype(np.int64))
        din = np.bincount(g.col.astype(np.int64), minlength=g.n)
        st = oracle.stats(g)
        assert st == {"n": g.n, "m": g.m, "max_out": int(dout.max()), "sum_out_sq": int((dout * dout).sum()),
                      "sum_in_out": int((din * dout).sum()), "rows_ge2": int((dout >= 2).sum())}


def _support_by_common_neighbours(sym, g):
    """support of every oriented arc (u, w) = |N(u) n N(w)| in the symmetric graph (textbook identity)."""
    nbr = [set(sym.col[sym.off[v]:sym.off[v + 1]].tolist()) for v in range(sym.n)]
    src = np.repeat(np.arange(g.n), np.diff(g.off.astype(np.int64)))
    return np.array([len(nbr[u] & nbr[w]) for u, w in zip(src.tolist(), g.col.tolist())], np.uint32), nbr


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_triangle_support_matches_common_neighbours_and_networkx(oracle, seed):
    n, s, d = G.random_multigraph(300, 2500, seed=seed)
    sym, g = oracle.clean_orient(n, s, d)
    sup, pv = oracle.triangle_support(g)
    want, nbr = _support_by_common_neighbours(sym, g)
    assert np.array_equal(sup, want)
    nxg = nx.Graph()
    nxg.add_nodes_from(range(sym.n))
    nxg.add_edges_from((u, w) for u in range(sym.n) for w in nbr[u] if u < w)
    tri = nx.triangles(nxg)
    assert pv.tolist() == [tri[v] for v in range(sym.n)]
    total = oracle.count(g)
    assert int(sup.sum()) == 3 * total and int(pv.sum()) == 3 * total


@pytest.mark.parametrize("k", [3, 4, 9, 30])
def test_triangle_support_closed_forms(oracle, k):
    _, g = oracle.clean_orient(*G.complete(k))
    sup, pv = oracle.triangle_support(g)
    assert set(sup.tolist()) == {k - 2} and set(pv.tolist()) == {(k - 1) * (k - 2) // 2}
    # wheel with r rim vertices: spokes carry 2 triangles, rim edges 1; the hub is in r, rim vertices in 2
    r = max(k, 4)
    sym, g = oracle.clean_orient(*G.wheel(r))
    sup, pv = oracle.triangle_support(g)
    deg = np.diff(sym.off.astype(np.int64))
    hub = int(np.argmax(deg))
    src = np.repeat(np.arange(g.n), np.diff(g.off.astype(np.int64)))
    spoke = (src == hub) | (g.col == hub)
    assert set(sup[spoke].tolist()) == {2} and set(sup[~spoke].tolist()) == {1}
    assert pv[hub] == r and set(np.delete(pv, hub).tolist()) == {2}
