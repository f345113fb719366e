"""GPU tests of the work and timing diagnostics: lstc_work_stats counts on the
device exactly what a host restatement of the planner's rules counts, and the
kernel timer sees every count launch."""
import numpy as np
import pytest

import graphs as G

pytestmark = pytest.mark.gpu


def work_model(off, col, in_off, in_col, n_owned, tiny_max, seg_words):
    """Host restatement of the plan's items and their reads (count_engine.hip k_plan_rows / k_count)."""
    off = off.astype(np.int64)
    n = len(off) - 1
    d = np.diff(off)
    split = in_off is not None
    rows = n if split else n_owned
    r = np.arange(rows)
    K = d[:rows]
    owned = r < n_owned
    kb = np.diff(in_off.astype(np.int64))[:rows] if split else np.zeros(rows, np.int64)
    active = (K >= 2) & (owned | (kb > 0))
    tiny = active & split & (K <= tiny_max)
    hashr = active & ~tiny
    nb = np.where(hashr, (kb + 511) // 512, 0)
    items = np.where(hashr, owned.astype(np.int64) + nb, 0)
    hashed = int((K * items).sum() + K[tiny].sum())
    src = np.repeat(np.arange(n), d)
    out_arcs = (src < rows)
    out_arcs[out_arcs] &= active[src[out_arcs]] & owned[src[out_arcs]]
    dw = d[col.astype(np.int64)]
    keep = out_arcs & (dw < d[src]) if split else out_arcs
    streamed = int(dw[keep].sum())
    segs_row = np.where(owned, K, 0) * active + kb * active
    if split:
        isrc = np.repeat(np.arange(n), np.diff(in_off.astype(np.int64)))
        sel = (isrc < rows)
        sel[sel] &= active[isrc[sel]]
        streamed += int(d[in_col.astype(np.int64)[sel]].sum())
    segments = int(segs_row[hashr].sum())
    tiny_segments = int(segs_row[tiny].sum())
    st = {"rows": rows, "items": int(items.sum()), "chunk_items": int(nb.sum()), "tiny_rows": int(tiny.sum()),
          "hashed_keys": hashed, "segments": segments, "tiny_segments": tiny_segments, "streamed": streamed}
    seg_bytes = 8 if (split and seg_words) else 20    # a word, or an id and two row offsets
    tiny_bytes = 12 if (split and seg_words) else 20  # an id and its word, or an id and two offsets
    st["bytes"] = (4 * (hashed + streamed) + seg_bytes * segments + tiny_bytes * tiny_segments
                   + 16 * (st["items"] + st["chunk_items"]) + 32 * st["tiny_rows"]
                   + 16 * (st["items"] + st["tiny_rows"]) + 2 * rows * (32 if split else 16))
    return st


def _graphs(lstc, oracle):
    out = []
    for n, s, d in [oracle.gen_rmat(13, 16, seed=4), oracle.gen_er(5000, 12, seed=5), G.complete(150),
                    oracle.gen_chung_lu(30000, 2.1, 30, 3000, seed=2)]:
        _, g = oracle.clean_orient(n, s, d)
        out.append(lstc.Graph.from_numpy(g.off, g.col))
    return out


def test_work_stats_match_host_model(lstc, oracle):
    gs = _graphs(lstc, oracle)
    plain = lstc.work_stats(gs)
    for g, st in zip(gs, plain):
        off, col = g.to_numpy()
        assert st == work_model(off, col, None, None, g.owned, 8, False)
    for g in gs:
        lstc.in_lists(g)  # owned sources; with segment words
    for tiny in (8, 4, 1):
        t = lstc.Tune(tiny_max=tiny)
        split = lstc.work_stats(gs, tune=t)
        for g, st in zip(gs, split):
            off, col = g.to_numpy()
            in_off = g.in_off[: g.n + 1].cpu().numpy().view(np.uint64)
            in_col = g.in_col[: int(in_off[-1])].cpu().numpy().view(np.uint32)
            assert st == work_model(off, col, in_off, in_col, g.owned, tiny if tiny > 1 else 0, True)
    # split counting streams less than the plain walk over the same arcs
    assert sum(s["streamed"] for s in split) < sum(s["streamed"] for s in plain)


def test_work_stats_of_halo_part_counts_owned_rows(lstc, oracle):
    n, s, d = oracle.gen_rmat(12, 16, seed=9)
    sym, _ = oracle.clean_orient(n, s, d)
    part = oracle.partition_random(sym.n, 4, seed=2)
    h = oracle.halo_extract(sym, part, 1, mode=0)
    g = lstc.Graph.from_numpy(h.csr.off, h.csr.col, n_owned=h.n_owned)
    st = lstc.work_stats([g])[0]
    assert st == work_model(h.csr.off, h.csr.col, None, None, h.n_owned, 8, False)
    lstc.in_lists(g)
    st = lstc.work_stats([g])[0]
    in_off = g.in_off[: g.n + 1].cpu().numpy().view(np.uint64)
    in_col = g.in_col[: int(in_off[-1])].cpu().numpy().view(np.uint32)
    assert st == work_model(h.csr.off, h.csr.col, in_off, in_col, h.n_owned, 8, True)


def test_kernel_timer_sees_every_launch(lstc, oracle):
    gs = _graphs(lstc, oracle)
    with lstc.KernelTimer() as t:
        for g in gs:
            lstc.count(g)
        lstc.count_batch(gs)
    assert t.launches == len(gs) + 1 and t.ms > 0
    with lstc.KernelTimer() as t2:
        pass
    assert t2.launches == 0
    # launches outside a timer are not recorded
    lstc.count(gs[0])
    with lstc.KernelTimer() as t3:
        lstc.count(gs[1])
    assert t3.launches == 1
