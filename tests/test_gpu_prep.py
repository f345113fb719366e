"""GPU parity of the generators and of clean + orient (lstc_clean_orient).

Bit-exact: raw samples, symmetric CSR and oriented CSR must equal the oracle's.
"""
import numpy as np
import pytest
import torch

import graphs as G

pytestmark = pytest.mark.gpu


def _np(t, dtype):
    return t.cpu().numpy().view(dtype)


@pytest.mark.parametrize("kind", ["rmat", "er", "chung_lu", "grid"])
def test_generators_bit_exact(lstc, oracle, kind):
    if kind == "rmat":
        n, s, d = lstc.generate_rmat(13, 16, seed=7)
        n2, s2, d2 = oracle.gen_rmat(13, 16, seed=7)
    elif kind == "er":
        n, s, d = lstc.generate_er(12345, 20, seed=3)
        n2, s2, d2 = oracle.gen_er(12345, 20, seed=3)
    elif kind == "chung_lu":
        n, s, d = lstc.generate_chung_lu(30000, 2.1, 50, 5e3, seed=5)
        n2, s2, d2 = oracle.gen_chung_lu(30000, 2.1, 50, 5e3, seed=5)
    else:
        n, s, d = lstc.generate_grid(37, 53, 0.3, seed=2)
        n2, s2, d2 = oracle.gen_grid(37, 53, 0.3, seed=2)
    assert n == n2
    assert np.array_equal(_np(s, np.uint32), s2)
    assert np.array_equal(_np(d, np.uint32), d2)


def test_partition_bit_exact(lstc, oracle):
    p = lstc.partition_random(100003, 8, seed=11)
    assert np.array_equal(_np(p, np.uint32), oracle.partition_random(100003, 8, seed=11))


def _check_clean_orient(lstc, oracle, n, s, d, compact=True):
    s = np.ascontiguousarray(s, np.uint32)
    d = np.ascontiguousarray(d, np.uint32)
    ds = torch.from_numpy(s.view(np.int32)).cuda() if len(s) else torch.zeros(0, dtype=torch.int32).cuda()
    dd = torch.from_numpy(d.view(np.int32)).cuda() if len(d) else torch.zeros(0, dtype=torch.int32).cuda()
    g = lstc.clean_orient(n, ds, dd, compact=compact, keep_symmetric=True)
    sym, og = oracle.clean_orient(n, s, d, compact)
    assert g.n == og.n and g.m == og.m
    off, col = g.to_numpy()
    assert np.array_equal(off, og.off)
    assert np.array_equal(col, og.col)
    soff, scol = g.sym_to_numpy()
    assert np.array_equal(soff, sym.off)
    assert np.array_equal(scol, sym.col)
    return g, og


@pytest.mark.parametrize("name,builder", [
    ("toy_test_cu", lambda O: G.toy_test_cu()),
    ("toy_edge_filter_cu", lambda O: G.toy_edge_filter_cu()),
    ("multigraph", lambda O: G.random_multigraph(3000, 20000, seed=5)),
    ("K200", lambda O: G.complete(200)),
    ("grid", lambda O: O.gen_grid(40, 40, 0.3, 1)),
    ("rmat_s12", lambda O: O.gen_rmat(12, 16, seed=3)),
    ("er", lambda O: O.gen_er(70000, 8, seed=4)),
])
@pytest.mark.parametrize("compact", [True, False])
def test_clean_orient_bit_exact(lstc, oracle, name, builder, compact):
    _check_clean_orient(lstc, oracle, *builder(oracle), compact=compact)


def test_clean_orient_degenerate(lstc, oracle):
    _check_clean_orient(lstc, oracle, 7, np.array([], np.uint32), np.array([], np.uint32))
    _check_clean_orient(lstc, oracle, 7, np.array([], np.uint32), np.array([], np.uint32), compact=False)
    _check_clean_orient(lstc, oracle, 9, np.array([1, 4, 4], np.uint32), np.array([1, 4, 4], np.uint32))
    _check_clean_orient(lstc, oracle, 9, np.array([8, 3, 8, 3, 3], np.uint32), np.array([3, 8, 3, 8, 8], np.uint32))


def test_clean_orient_rejects_out_of_range(lstc):
    s = torch.tensor([0, 5], dtype=torch.int32).cuda()
    d = torch.tensor([1, 2], dtype=torch.int32).cuda()
    with pytest.raises(lstc.LstcError):
        lstc.clean_orient(4, s, d)


def test_device_pipeline_rmat_s16(lstc, oracle):
    """Whole device path at the CPU-runnable config: generate -> clean/orient -> count."""
    n, s, d = lstc.generate_rmat(16, 16, seed=1)
    g = lstc.clean_orient(n, s, d)
    sym, og = oracle.clean_orient(*oracle.gen_rmat(16, 16, seed=1))
    off, col = g.to_numpy()
    assert np.array_equal(off, og.off) and np.array_equal(col, og.col)
    assert lstc.count(g) == oracle.count(og)


@pytest.mark.parametrize("name,builder", [
    ("toy_edge_filter_cu", lambda O: G.toy_edge_filter_cu()),
    ("K120", lambda O: G.complete(120)),
    ("rmat_s13", lambda O: O.gen_rmat(13, 16, seed=5)),
    ("er", lambda O: O.gen_er(30000, 12, seed=6)),
    ("grid", lambda O: O.gen_grid(30, 40, 0.3, 2)),
])
def test_in_lists_bit_exact(lstc, oracle, name, builder):
    n, s, d = builder(oracle)
    ds = torch.from_numpy(np.ascontiguousarray(s, np.uint32).view(np.int32)).cuda()
    dd = torch.from_numpy(np.ascontiguousarray(d, np.uint32).view(np.int32)).cuda()
    g = lstc.clean_orient(n, ds, dd, in_lists=True)
    _, og = oracle.clean_orient(n, s, d)
    want_off, want_col = G.owned_in_lists(og.off, og.col, og.n)
    got_off = g.in_off[: g.n + 1].cpu().numpy().view(np.uint64)
    got_col = g.in_col[: int(want_off[-1])].cpu().numpy().view(np.uint32)
    assert np.array_equal(got_off, want_off)
    assert np.array_equal(got_col, want_col)
    assert g.meta["in_entries"] == int(want_off[-1])
    # segment words written by the same pass (lstc_clean_orient_split): start | length << 32 of each entry's row
    deg = np.diff(og.off.astype(np.int64)).astype(np.uint64)
    want_cs = (og.off[:-1] & np.uint64(0xFFFFFFFF)) | (deg << np.uint64(32))
    assert np.array_equal(g.col_seg[: g.m].cpu().numpy().view(np.uint64), want_cs[og.col])
    assert np.array_equal(g.in_seg[: len(want_col)].cpu().numpy().view(np.uint64), want_cs[want_col])
    # split count (each arc streams its shorter list) equals the oracle
    assert lstc.count(g) == oracle.count(og)
    assert lstc.count(g, split=False) == oracle.count(og)


@pytest.mark.parametrize("name,n,s,d", [
    ("no_edges", 7, [], []),
    ("loops_only", 5, [1, 2, 3], [1, 2, 3]),                      # every pair dropped: no symmetric arcs
    ("one_edge", 9, [3], [7]),
    ("duplicates", 9, [8, 3, 8, 3, 3], [3, 8, 3, 8, 8]),
    ("triangle_and_isolated", 12, [0, 1, 2, 2], [1, 2, 0, 2]),
])
def test_split_preprocessing_degenerate(lstc, oracle, name, n, s, d):
    """lstc_clean_orient_split (in-lists and segment words through the chained scan) on inputs
    with no, one or only duplicate / looping edges: the same CSR and in-lists as the oracle's."""
    s = np.array(s, np.uint32)
    d = np.array(d, np.uint32)
    ds = torch.from_numpy(s.view(np.int32)).cuda() if len(s) else torch.zeros(0, dtype=torch.int32).cuda()
    dd = torch.from_numpy(d.view(np.int32)).cuda() if len(d) else torch.zeros(0, dtype=torch.int32).cuda()
    g = lstc.clean_orient(n, ds, dd, in_lists=True)
    _, og = oracle.clean_orient(n, s, d)
    assert g.n == og.n and g.m == og.m
    off, col = g.to_numpy()
    assert np.array_equal(off, og.off) and np.array_equal(col, og.col)
    want_off, want_col = G.owned_in_lists(og.off, og.col, og.n)
    assert g.meta["in_entries"] == int(want_off[-1])
    assert np.array_equal(g.in_off[: g.n + 1].cpu().numpy().view(np.uint64), want_off)
    if len(want_col):
        assert np.array_equal(g.in_col[: len(want_col)].cpu().numpy().view(np.uint32), want_col)
    assert lstc.count(g) == oracle.count(og)


def test_split_preprocessing_many_tiles(lstc, oracle):
    """A graph whose unique keys span hundreds of 4096-key tiles, so the in-list positions come
    through long look-back chains: in-lists and segment words equal the oracle's."""
    n, s, d = oracle.gen_rmat(17, 16, seed=12)
    ds = torch.from_numpy(np.ascontiguousarray(s, np.uint32).view(np.int32)).cuda()
    dd = torch.from_numpy(np.ascontiguousarray(d, np.uint32).view(np.int32)).cuda()
    g = lstc.clean_orient(n, ds, dd, in_lists=True)
    _, og = oracle.clean_orient(n, s, d)
    want_off, want_col = G.owned_in_lists(og.off, og.col, og.n)
    assert np.array_equal(g.in_off[: g.n + 1].cpu().numpy().view(np.uint64), want_off)
    assert np.array_equal(g.in_col[: len(want_col)].cpu().numpy().view(np.uint32), want_col)
    deg = np.diff(og.off.astype(np.int64)).astype(np.uint64)
    want_cs = (og.off[:-1] & np.uint64(0xFFFFFFFF)) | (deg << np.uint64(32))
    assert np.array_equal(g.col_seg[: g.m].cpu().numpy().view(np.uint64), want_cs[og.col])
    assert np.array_equal(g.in_seg[: len(want_col)].cpu().numpy().view(np.uint64), want_cs[want_col])
    assert lstc.count(g) == oracle.count(og)


def _stream_copy_expected(og):
    """Row starts of the streaming copy: d+ rounded up to 32 entries, scanned."""
    deg = np.diff(og.off.astype(np.int64))
    pad = (deg + 31) // 32 * 32
    start = np.concatenate([[0], np.cumsum(pad)]).astype(np.int64)
    return deg, start


@pytest.mark.parametrize("name,builder", [
    ("toy_edge_filter_cu", lambda O: G.toy_edge_filter_cu()),
    ("K120", lambda O: G.complete(120)),
    ("rmat_s13", lambda O: O.gen_rmat(13, 16, seed=5)),
    ("er", lambda O: O.gen_er(30000, 12, seed=6)),
    ("grid", lambda O: O.gen_grid(30, 40, 0.3, 2)),
    ("rmat_s17", lambda O: O.gen_rmat(17, 16, seed=12)),
    ("no_edges", lambda O: (7, np.zeros(0, np.uint32), np.zeros(0, np.uint32))),
    ("one_edge", lambda O: (9, np.array([3], np.uint32), np.array([7], np.uint32))),
])
def test_stream_copy_bit_exact(lstc, oracle, name, builder):
    """lstc_clean_orient_stream: the same CSR and in-lists as the oracle's, every row again in the
    streaming copy at a 32-entry aligned start (d+ rounded up to 32, scanned), the segment words
    holding those starts; the count that streams from the copy (lstc_count_split_stream), the
    batched count (which walks without the words) and the count after the in-lists are rebuilt
    all equal the oracle's."""
    n, s, d = builder(oracle)
    s = np.ascontiguousarray(s, np.uint32)
    d = np.ascontiguousarray(d, np.uint32)
    ds = torch.from_numpy(s.view(np.int32)).cuda() if len(s) else torch.zeros(0, dtype=torch.int32).cuda()
    dd = torch.from_numpy(d.view(np.int32)).cuda() if len(d) else torch.zeros(0, dtype=torch.int32).cuda()
    g = lstc.clean_orient(n, ds, dd, in_lists=True, stream_copy=True)
    _, og = oracle.clean_orient(n, s, d)
    assert g.n == og.n and g.m == og.m
    off, col = g.to_numpy()
    assert np.array_equal(off, og.off) and np.array_equal(col, og.col)
    want_off, want_col = G.owned_in_lists(og.off, og.col, og.n)
    assert np.array_equal(g.in_off[: g.n + 1].cpu().numpy().view(np.uint64), want_off)
    assert np.array_equal(g.in_col[: len(want_col)].cpu().numpy().view(np.uint32), want_col)
    deg, start = _stream_copy_expected(og)
    assert int(g.scol.numel()) == max(int(start[-1]), 1)
    scol = g.scol.cpu().numpy().view(np.uint32)
    for v in np.flatnonzero(deg):
        assert np.array_equal(scol[start[v]: start[v] + deg[v]], og.col[og.off[v]: og.off[v + 1]]), (name, v)
    words = (start[:-1].astype(np.uint64) & np.uint64(0xFFFFFFFF)) | (deg.astype(np.uint64) << np.uint64(32))
    assert np.array_equal(g.col_seg[: g.m].cpu().numpy().view(np.uint64), words[og.col])
    assert np.array_equal(g.in_seg[: len(want_col)].cpu().numpy().view(np.uint64), words[want_col])
    want = oracle.count(og)
    assert lstc.count(g) == want
    for t in (lstc.Tune(unroll=8), lstc.Tune(waves=2), lstc.Tune(waves=8), lstc.Tune(tiny_max=1, table_scale=2)):
        assert lstc.count(g, tune=t) == want, (name, t.as_dict())
    assert lstc.count_batch([g]) == [want]
    lstc.in_lists(g)  # words into col again: the copy is dropped
    assert g.scol is None and lstc.count(g) == want
