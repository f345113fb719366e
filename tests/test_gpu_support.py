"""GPU parity of per-edge triangle support and per-vertex triangle counts
(lstc_triangle_support) against the oracle's lo_triangle_support: bit-exact per arc
and per vertex, including rows longer than one LDS tile and halo parts (owned rows)."""
import numpy as np
import pytest

import graphs as G

pytestmark = pytest.mark.gpu


def _check(lstc, oracle, og, n_owned=None):
    g = lstc.Graph.from_numpy(og.off, og.col, n_owned=n_owned)
    sup, pv = lstc.triangle_support(g)
    want_sup, want_pv = oracle.triangle_support(og, n_owned)
    assert np.array_equal(sup.cpu().numpy().view(np.uint32), want_sup)
    assert np.array_equal(pv.cpu().numpy().view(np.uint64), want_pv)
    total = lstc.count(g)
    assert int(want_sup.sum()) == 3 * total
    # each output alone
    s2, none = lstc.triangle_support(g, per_vertex=False)
    assert none is None and np.array_equal(s2.cpu().numpy().view(np.uint32), want_sup)
    none, p2 = lstc.triangle_support(g, per_edge=False)
    assert none is None and np.array_equal(p2.cpu().numpy().view(np.uint64), want_pv)
    # split rule (lstc_triangle_support_split): owned-source in-lists, every arc streams its shorter list
    lstc.in_lists(g)
    assert g.split_ready()
    for pe, pvv in [(True, True), (True, False), (False, True)]:
        s3, p3 = lstc.triangle_support(g, per_edge=pe, per_vertex=pvv, split=True)
        if pe:
            assert np.array_equal(s3.cpu().numpy().view(np.uint32), want_sup)
        if pvv:
            assert np.array_equal(p3.cpu().numpy().view(np.uint64), want_pv)


@pytest.mark.parametrize("name,builder", [
    ("K3", lambda O: G.complete(3)),
    ("K200", lambda O: G.complete(200)),
    ("K1500_tiled", lambda O: G.complete(1500)),
    ("wheel", lambda O: G.wheel(50)),
    ("rmat_s13", lambda O: O.gen_rmat(13, 16, seed=3)),
    ("er", lambda O: O.gen_er(20000, 16, seed=4)),
    ("grid", lambda O: O.gen_grid(100, 80, 0.3, 2)),
    ("chung_lu", lambda O: O.gen_chung_lu(30000, 2.1, 30, 4000, seed=5)),
    ("chung_lu_hubs", lambda O: O.gen_chung_lu(60000, 2.1, 40, 2e4, seed=7)),  # in-lists over 512 sources
    ("rmat_s15", lambda O: O.gen_rmat(15, 16, seed=8)),
    ("star", lambda O: G.star(500)),
])
def test_support_matches_oracle(lstc, oracle, name, builder):
    _, og = oracle.clean_orient(*builder(oracle))
    _check(lstc, oracle, og)


def test_support_of_halo_parts(lstc, oracle):
    n, s, d = oracle.gen_rmat(12, 16, seed=6)
    sym, _ = oracle.clean_orient(n, s, d)
    part = oracle.partition_random(sym.n, 3, seed=2)
    for p in range(3):
        h = oracle.halo_extract(sym, part, p, mode=0)
        _check(lstc, oracle, h.csr, h.n_owned)


def test_support_empty_and_invalid(lstc, oracle):
    _, og = oracle.clean_orient(*G.star(10))
    g = lstc.Graph.from_numpy(og.off, og.col)
    sup, pv = lstc.triangle_support(g)
    assert int(sup.sum().item()) == 0 and int(pv.sum().item()) == 0
    with pytest.raises(lstc.LstcError):
        lstc.triangle_support(g, per_edge=False, per_vertex=False)


def test_support_tiled_workgroup_rows(lstc, oracle):
    """K_2500: rows up to d+ = 2499 are hashed in two workgroup tiles of 2048 keys.  Closed
    forms: every arc's support is n - 2, every vertex's count (n - 1)(n - 2) / 2, on the plain
    walk and on the split rule."""
    n = 2500
    _, og = oracle.clean_orient(*G.complete(n))
    assert int(np.diff(og.off).max()) == n - 1
    g = lstc.Graph.from_numpy(og.off, og.col)
    for split in (False, True):
        if split:
            lstc.in_lists(g)
        sup, pv = lstc.triangle_support(g, split=split)
        s = sup.cpu().numpy().view(np.uint32)
        p = pv.cpu().numpy().view(np.uint64)
        assert s.shape == (n * (n - 1) // 2,) and bool((s == n - 2).all()), split
        assert bool((p == (n - 1) * (n - 2) // 2).all()), split
