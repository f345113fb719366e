"""GPU parity of the counting kernel (lstc_count) against the CPU oracle.

Every GPU result must equal the oracle's uint64 exactly.
"""
import copy

import numpy as np
import pytest
import torch

import graphs as G

pytestmark = pytest.mark.gpu


def _oracle_graph(O, builder):
    n, s, d = builder
    sym, g = O.clean_orient(n, s, d)
    return sym, g


def _upload(T, g, sym=None):
    if sym is None:
        return T.Graph.from_numpy(g.off, g.col)
    return T.Graph.from_numpy(g.off, g.col, sym.off, sym.col)


@pytest.mark.parametrize("name,builder", [
    ("K3", lambda O: G.complete(3)),
    ("K40", lambda O: G.complete(40)),
    ("K300", lambda O: G.complete(300)),
    ("K1200_tiled", lambda O: G.complete(1200)),
    ("wheel50", lambda O: G.wheel(50)),
    ("windmill300", lambda O: G.windmill(300)),
    ("star", lambda O: G.star(5000)),
    ("bipartite", lambda O: G.complete_bipartite(40, 70)),
    ("toy_test_cu", lambda O: G.toy_test_cu()),
    ("toy_edge_filter_cu", lambda O: G.toy_edge_filter_cu()),
    ("grid", lambda O: O.gen_grid(50, 70, 0.3, 5)),
    ("er", lambda O: O.gen_er(20000, 20, seed=3)),
    ("chung_lu", lambda O: O.gen_chung_lu(50000, 2.1, 50, 5e3, seed=2)),
    ("rmat_s10", lambda O: O.gen_rmat(10, 16, seed=1)),
    ("rmat_s14", lambda O: O.gen_rmat(14, 16, seed=2)),
])
def test_count_matches_oracle(lstc, oracle, name, builder):
    sym, g = _oracle_graph(oracle, builder(oracle))
    want = oracle.count(g)
    got = lstc.count(_upload(lstc, g))
    assert got == want, name


def test_count_rmat_s16_config(lstc, oracle):
    """RMAT scale 16, ef 16, a/b/c/d .57/.19/.19/.05, seed 1 (the CPU-runnable case)."""
    sym, g = _oracle_graph(oracle, oracle.gen_rmat(16, 16, seed=1))
    want = oracle.count(g)
    dg = _upload(lstc, g)
    assert lstc.count(dg) == want
    # repeatable bit-for-bit
    assert lstc.count(dg) == want


def test_per_row_matches_reference_partials(lstc, oracle):
    """d_per_row equals the reference's per-block partials g_sum[u] (K1 restatement)."""
    for builder in [oracle.gen_rmat(12, 16, seed=9), G.complete(700), oracle.gen_er(5000, 30, seed=1)]:
        _, g = _oracle_graph(oracle, builder)
        t_ref, rows_ref = oracle.count_k1(g)
        t, rows = lstc.count(_upload(lstc, g), per_row=True)
        assert t == t_ref
        assert np.array_equal(rows, rows_ref)


def test_owned_rows(lstc, oracle):
    _, g = _oracle_graph(oracle, oracle.gen_rmat(13, 16, seed=4))
    dg = _upload(lstc, g)
    for cut in [0, 1, 777, g.n // 2, g.n]:
        assert lstc.count(dg, n_owned=cut) == oracle.count(g, n_owned=cut), cut


def test_empty_and_edgeless(lstc, oracle):
    g0 = lstc.Graph.from_numpy(np.zeros(1, np.uint64), np.zeros(0, np.uint32))
    assert lstc.count(g0) == 0
    g1 = lstc.Graph.from_numpy(np.zeros(11, np.uint64), np.zeros(0, np.uint32))
    assert lstc.count(g1) == 0


def test_invalid_arguments_raise(lstc, oracle):
    _, g = _oracle_graph(oracle, G.complete(10))
    dg = _upload(lstc, g)
    with pytest.raises(lstc.LstcError):
        lstc.count(dg, n_owned=dg.n + 1)


@pytest.mark.parametrize("name,builder", [
    ("K3", lambda O: G.complete(3)),
    ("K1200", lambda O: G.complete(1200)),
    ("K3000_tiled", lambda O: G.complete(3000)),
    ("wheel", lambda O: G.wheel(300)),
    ("star", lambda O: G.star(9000)),
    ("chung_lu", lambda O: O.gen_chung_lu(60000, 2.1, 50, 2e4, seed=4)),
    ("rmat_s15", lambda O: O.gen_rmat(15, 16, seed=8)),
    ("grid", lambda O: O.gen_grid(300, 200, 0.3, 3)),
    ("er", lambda O: O.gen_er(30000, 12, seed=9)),
    ("toy_edge_filter_cu", lambda O: G.toy_edge_filter_cu()),
])
def test_split_count_matches_oracle(lstc, oracle, name, builder):
    n, s, d = builder(oracle)
    ds = torch.from_numpy(np.ascontiguousarray(s, np.uint32).view(np.int32)).cuda()
    dd = torch.from_numpy(np.ascontiguousarray(d, np.uint32).view(np.int32)).cuda()
    g = lstc.clean_orient(n, ds, dd, in_lists=True)
    _, og = oracle.clean_orient(n, s, d)
    # complete graphs: closed form n(n-1)(n-2)/6 (the merge oracle is slow at K3000)
    want = n * (n - 1) * (n - 2) // 6 if name.startswith("K") else oracle.count(og)
    assert lstc.count(g) == want, name
    for t in (lstc.Tune(table_scale=2), lstc.Tune(table_scale=4, small_max=64), lstc.Tune(unroll=8),
              lstc.Tune(tiny_max=1), lstc.Tune(tiny_max=2), lstc.Tune(tiny_max=5, small_group=4),
              lstc.Tune(table_scale=16, small_max=32),
              # workgroup size 128 / 512: big rows tile by 1024 / 4096 keys
              lstc.Tune(waves=2), lstc.Tune(waves=8), lstc.Tune(waves=2, table_scale=2, unroll=8),
              lstc.Tune(waves=8, small_max=64, grid_per_cu=1)):
        assert lstc.count(g, tune=t) == want, (name, t.as_dict())
    # a batch mixing split and plain graphs
    plain = lstc.Graph.from_numpy(og.off, og.col)
    assert lstc.count_batch([g, plain, g]) == [want, want, want]
    # without segment words each segment's list is located through the row offsets
    g2 = copy.copy(g)
    g2.col_seg = g2.in_seg = None
    assert lstc.count(g2) == want, name
    assert lstc.count(g2, tune=lstc.Tune(table_scale=2, unroll=8)) == want, name
    assert lstc.count_batch([g2, g, plain]) == [want, want, want]


@pytest.mark.parametrize("cut_frac", [0.0, 0.3, 1.0])
def test_in_lists_of_any_csr(lstc, oracle, cut_frac):
    """lstc_in_lists_segments (and lstc_in_lists + lstc_segments) on an uploaded CSR with n_owned
    rows: definition, segment words of every entry, and split = plain count."""
    _, og = oracle.clean_orient(*oracle.gen_rmat(13, 16, seed=12))
    cut = int(og.n * cut_frac)
    g = lstc.Graph.from_numpy(og.off, og.col, n_owned=cut)
    lstc.in_lists(g)
    want_off, want_col = G.owned_in_lists(og.off, og.col, cut)
    assert np.array_equal(g.in_off.cpu().numpy().view(np.uint64), want_off)
    assert np.array_equal(g.in_col[: len(want_col)].cpu().numpy().view(np.uint32), want_col)
    want = oracle.count(og, n_owned=cut)
    assert lstc.count(g) == want
    # in-lists built for another n_owned are not used
    assert lstc.count(g, n_owned=og.n) == oracle.count(og)
    # segment words: start | length << 32 of every col entry's and in-list entry's row
    d = np.diff(og.off.astype(np.int64))
    def words(ids):
        ids = ids.astype(np.int64)
        return (og.off[ids].astype(np.uint64) & np.uint64(0xFFFFFFFF)) | (d[ids].astype(np.uint64) << np.uint64(32))
    # the fused build writes the column words of owned rows (the ones the count reads)
    owned_arcs = int(og.off[cut])
    assert np.array_equal(g.col_seg[:owned_arcs].cpu().numpy().view(np.uint64), words(og.col[:owned_arcs]))
    assert np.array_equal(g.in_seg[: len(want_col)].cpu().numpy().view(np.uint64), words(want_col))
    # lstc_in_lists alone followed by lstc_segments: the same in-lists, words of every entry
    g2 = lstc.Graph.from_numpy(og.off, og.col, n_owned=cut)
    lstc.in_lists(g2, segment_words=False)
    assert g2.col_seg is None and lstc.count(g2) == want
    lstc.segments(g2)
    assert np.array_equal(g2.col_seg[: og.m].cpu().numpy().view(np.uint64), words(og.col))
    for a, b in ((g.in_off, g2.in_off), (g.in_col[: len(want_col)], g2.in_col[: len(want_col)]),
                 (g.col_seg[:owned_arcs], g2.col_seg[:owned_arcs]), (g.in_seg[: len(want_col)], g2.in_seg[: len(want_col)])):
        assert torch.equal(a, b)
    assert lstc.count(g2) == want


def test_count_above_2_30_arcs(lstc):
    """m > 2^30 arcs, so column byte offsets need more than 32 bits: disjoint K64
    copies with id-ordered arcs, built in HBM directly.  Closed form: copies * C(64, 3),
    plain and split (in-lists and segment words); per-edge support k - 2 on every arc and
    per-vertex counts C(k - 1, 2) on every vertex, from both support walks."""
    k = 64
    per = k * (k - 1) // 2
    copies = (1 << 30) // per + 1
    n, m = k * copies, per * copies
    dev = torch.device("cuda")
    d = k - 1 - torch.arange(k, dtype=torch.int64)
    row0 = torch.cumsum(d, 0) - d  # start of local row i inside a clique's block
    tcol = torch.cat([torch.arange(i + 1, k, dtype=torch.int32) for i in range(k)]).to(dev)
    base = torch.arange(copies, dtype=torch.int32, device=dev) * k
    col = (tcol[None, :] + base[:, None]).reshape(-1)
    off = torch.empty(n + 1, dtype=torch.int64, device=dev)
    off[:n] = (torch.arange(copies, dtype=torch.int64, device=dev)[:, None] * per + row0.to(dev)[None, :]).reshape(-1)
    off[n] = m
    del base
    g = lstc.Graph(n, m, off, col)
    want = copies * (k * (k - 1) * (k - 2) // 6)
    assert m > (1 << 30)
    assert lstc.count(g) == want
    lstc.in_lists(g)
    assert lstc.count(g) == want
    assert lstc.count(g, tune=lstc.Tune(unroll=8)) == want
    for split in (False, True):
        sup, pv = lstc.triangle_support(g, split=split)
        assert sup.shape[0] == m and bool((sup == k - 2).all()), split
        assert pv.shape[0] == n and bool((pv == (k - 1) * (k - 2) // 2).all()), split
        del sup, pv
