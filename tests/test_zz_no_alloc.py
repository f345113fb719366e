"""GPU test of lstc_count_split_static, the entry point that allocates nothing: called
through the C ABI, it must give exactly what lstc_count_split gives (and the oracle), on
consecutive calls whose graphs shrink and grow again, and refuse graphs whose scratch need
exceeds its static array."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _split_graph(lstc, oracle, gen):
    n, s, d = gen
    _, g = oracle.clean_orient(n, s, d)
    dg = lstc.Graph.from_numpy(g.off, g.col)
    lstc.in_lists(dg)  # owned-source in-lists and segment words
    return dg, oracle.count(g)


def _count_static(lstc, g, out, tune=None):
    L = lstc.load()
    st = L.lstc_count_split_static(_p(g.off), _p(g.col), _p(g.in_off), _p(g.in_col), _p(g.col_seg), _p(g.in_seg),
                                   g.n, g.owned, g.m, _p(out), C.byref(tune) if tune is not None else None,
                                   C.c_void_p(lstc.stream_handle()))
    lstc.check(st, "lstc_count_split_static")
    torch.cuda.synchronize()
    return int(out[0].item())


def test_static_matches_workspace_entry_point(lstc, oracle):
    big, want_big = _split_graph(lstc, oracle, oracle.gen_rmat(15, 16, seed=21))
    small, want_small = _split_graph(lstc, oracle, oracle.gen_er(3000, 12, seed=4))
    L = lstc.load()
    assert L.lstc_count_split_static_limit() == 32 << 20
    out = torch.full((1,), -1, dtype=torch.int64, device="cuda")
    ref_big, ref_small = lstc.count(big), lstc.count(small)  # lstc_count_split with a workspace
    assert (ref_big, ref_small) == (want_big, want_small)
    # two calls in a row, the second smaller; then larger again, and with another tune
    assert _count_static(lstc, big, out) == ref_big
    assert _count_static(lstc, small, out) == ref_small
    assert _count_static(lstc, big, out) == ref_big
    assert _count_static(lstc, small, out, lstc.Tune(table_scale=4, unroll=8, tiny_max=1)) == ref_small


def test_static_refuses_oversized_graphs(lstc, oracle):
    g, _ = _split_graph(lstc, oracle, oracle.gen_rmat(10, 8, seed=2))
    L = lstc.load()
    out = torch.full((1,), -1, dtype=torch.int64, device="cuda")
    need = C.c_size_t()
    lstc.check(L.lstc_count_workspace(1 << 23, 1 << 23, 1 << 20, None, C.byref(need)))
    assert need.value > L.lstc_count_split_static_limit()
    # row count far beyond the scratch (the call is refused before any launch)
    st = L.lstc_count_split_static(_p(g.off), _p(g.col), _p(g.in_off), _p(g.in_col), None, None, 1 << 23, 1 << 23,
                                   1 << 20, _p(out), None, C.c_void_p(lstc.stream_handle()))
    assert st == -1 and b"static scratch" in L.lstc_last_error()
    torch.cuda.synchronize()
    assert int(out[0].item()) == -1
