"""GPU parity of halo extraction (lstc_halo_sizes / lstc_halo_extract).

Bit-exact CSR equality with the oracle's restatement of round_clean_subgraph.cu,
and the invariant that the parts' owned-row counts sum to the whole count.
"""
import numpy as np
import pytest
import torch

import graphs as G

pytestmark = pytest.mark.gpu


def _device_graph(lstc, n, s, d):
    ds = torch.from_numpy(np.ascontiguousarray(s, np.uint32).view(np.int32)).cuda()
    dd = torch.from_numpy(np.ascontiguousarray(d, np.uint32).view(np.int32)).cuda()
    return lstc.clean_orient(n, ds, dd, keep_symmetric=True)


@pytest.mark.parametrize("name,builder,k", [
    ("multigraph", lambda O: G.random_multigraph(2000, 12000, seed=3), 3),
    ("rmat_s12", lambda O: O.gen_rmat(12, 16, seed=2), 8),
    ("grid", lambda O: O.gen_grid(30, 30, 0.3, 4), 4),
    ("K100", lambda O: G.complete(100), 2),
])
@pytest.mark.parametrize("mode", [0, 1])
def test_halo_bit_exact(lstc, oracle, name, builder, k, mode):
    n, s, d = builder(oracle)
    g = _device_graph(lstc, n, s, d)
    sym, og = oracle.clean_orient(n, s, d)
    part = lstc.partition_random(g.n, k, seed=5)
    opart = oracle.partition_random(sym.n, k, seed=5)
    for p in range(k):
        h = lstc.halo_extract(g, part, p, mode=mode, with_l2g=True)
        oh = oracle.halo_extract(sym, opart, p, mode)
        assert h.n_owned == oh.n_owned and h.n == oh.csr.n and h.m == oh.csr.m, (name, p)
        off, col = h.to_numpy()
        assert np.array_equal(off, oh.csr.off), (name, p)
        assert np.array_equal(col, oh.csr.col), (name, p)
        assert np.array_equal(h.meta["l2g"].cpu().numpy().view(np.uint32), oh.l2g), (name, p)


@pytest.mark.parametrize("k,mode", [(2, 0), (8, 0), (8, 1)])
def test_halo_parts_sum_to_unpartitioned(lstc, oracle, k, mode):
    n, s, d = lstc.generate_rmat(15, 16, seed=3)
    g = lstc.clean_orient(n, s, d, keep_symmetric=True)
    total = lstc.count(g)
    part = lstc.partition_random(g.n, k, seed=1)
    acc = 0
    for p in range(k):
        h = lstc.halo_extract(g, part, p, mode=mode)
        acc += lstc.count(h)  # counts owned rows [0, n_owned)
    assert acc == total


def test_halo_empty_part(lstc, oracle):
    n, s, d = G.complete(20)
    g = _device_graph(lstc, n, s, d)
    part = torch.zeros(g.n, dtype=torch.int32, device="cuda")
    h = lstc.halo_extract(g, part, 1)  # nobody is in part 1
    assert h.n == 0 and h.m == 0 and h.n_owned == 0
    h = lstc.halo_extract(g, part, 0)
    assert h.n == 20 and h.n_owned == 20 and lstc.count(h) == 20 * 19 * 18 // 6


@pytest.mark.parametrize("name,builder,k", [
    ("rmat_s13", lambda O: O.gen_rmat(13, 16, seed=6), 8),
    ("grid", lambda O: O.gen_grid(40, 40, 0.3, 4), 4),
    ("chung_lu", lambda O: O.gen_chung_lu(20000, 2.1, 30, 4e3, seed=3), 3),
])
@pytest.mark.parametrize("mode", [0, 1])
def test_halo_split_parts_match_oracle(lstc, oracle, name, builder, k, mode):
    """Halo parts with owned-source in-lists (lstc_in_lists) counted split, part by part and batched."""
    n, s, d = builder(oracle)
    g = _device_graph(lstc, n, s, d)
    sym, og = oracle.clean_orient(n, s, d)
    whole = oracle.count(og)
    part = lstc.partition_random(g.n, k, seed=7)
    opart = oracle.partition_random(sym.n, k, seed=7)
    parts, acc = [], 0
    for p in range(k):
        h = lstc.halo_extract(g, part, p, mode=mode, split=True)
        oh = oracle.halo_extract(sym, opart, p, mode)
        want_off, want_col = G.owned_in_lists(oh.csr.off, oh.csr.col, oh.n_owned)
        assert np.array_equal(h.in_off.cpu().numpy().view(np.uint64), want_off), (name, p)
        assert h.meta["in_entries"] == len(want_col)
        assert np.array_equal(h.in_col[: len(want_col)].cpu().numpy().view(np.uint32), want_col), (name, p)
        want = oracle.count(oh.csr, n_owned=oh.n_owned)
        assert lstc.count(h) == want, (name, p)
        assert lstc.count(h, split=False) == want, (name, p)
        for t in (lstc.Tune(tiny_max=1), lstc.Tune(tiny_max=4), lstc.Tune(unroll=8, table_scale=2)):
            assert lstc.count(h, tune=t) == want, (name, p, t.as_dict())
        parts.append(h)
        acc += want
    assert acc == whole
    assert lstc.count_batch(parts) == [lstc.count(h) for h in parts]
