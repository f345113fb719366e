"""GPU parity of the batched path: one persistent launch over many graphs.

Each graph's count must be identical alone (lstc_count), inside a batch
(lstc_count_batch), through a resident plan, and through its hipGraph replay.
"""
import numpy as np
import pytest
import torch

import graphs as G

pytestmark = pytest.mark.gpu


def _mixed(lstc, oracle):
    """Uneven graphs: RMAT, ER of several sizes, a grid, a clique, an empty graph."""
    out, want = [], []
    specs = [oracle.gen_rmat(12, 16, seed=1), oracle.gen_er(3000, 20, seed=2), oracle.gen_er(500, 8, seed=3),
             oracle.gen_grid(40, 30, 0.3, 4), G.complete(90), oracle.gen_rmat(9, 4, seed=5),
             G.star(100), oracle.gen_chung_lu(20000, 2.1, 30, 2000, seed=6)]
    for n, s, d in specs:
        sym, g = oracle.clean_orient(n, s, d)
        out.append(lstc.Graph.from_numpy(g.off, g.col))
        want.append(oracle.count(g))
    return out, want


def test_batch_matches_oracle_and_single(lstc, oracle):
    gs, want = _mixed(lstc, oracle)
    assert lstc.count_batch(gs) == want
    assert [lstc.count(g) for g in gs] == want


def test_batch_of_one_and_duplicates(lstc, oracle):
    gs, want = _mixed(lstc, oracle)
    assert lstc.count_batch(gs[:1]) == want[:1]
    # the same graph many times in one launch
    assert lstc.count_batch([gs[0]] * 7 + [gs[4]] * 3) == [want[0]] * 7 + [want[4]] * 3


def test_batch_halo_parts_sum(lstc, oracle):
    """All k halo parts of a partition in one launch; the parts sum to the whole."""
    n, s, d = lstc.generate_rmat(15, 16, seed=7)
    g = lstc.clean_orient(n, s, d, keep_symmetric=True)
    total = lstc.count(g)
    part = lstc.partition_random(g.n, 8, seed=3)
    for mode in (0, 1):
        parts = [lstc.halo_extract(g, part, p, mode=mode) for p in range(8)]
        per = lstc.count_batch(parts)
        assert per == [lstc.count(h) for h in parts]
        assert sum(per) == total


def test_plan_run_and_graph_replay(lstc, oracle):
    gs, want = _mixed(lstc, oracle)
    plan = lstc.BatchPlan(gs)
    try:
        plan.run()
        assert plan.results() == want
        for _ in range(3):  # replays, with eager work and synchronisations in between
            plan.counts.fill_(12345)
            plan.launch_graph()
            assert plan.results() == want
            assert lstc.count(gs[1]) == want[1]
    finally:
        plan.close()


def test_split_plan_run_and_graph_replay(lstc, oracle):
    """Graphs carrying in-lists and segment words (and plain ones) in a plan and its hipGraph replays."""
    gs, want = _mixed(lstc, oracle)
    for g in gs[::2]:
        lstc.in_lists(g)  # split counting with segment words for every other graph
    for g in gs[1:4:2]:
        lstc.in_lists(g)
        g.col_seg = g.in_seg = None  # split counting through the row offsets
    assert lstc.count_batch(gs) == want
    plan = lstc.BatchPlan(gs, tune=lstc.Tune(table_scale=4, unroll=8))
    try:
        plan.run()
        assert plan.results() == want
        for _ in range(3):
            plan.counts.fill_(777)
            plan.launch_graph()
            assert plan.results() == want
            assert lstc.count(gs[0]) == want[0]
    finally:
        plan.close()


@pytest.mark.parametrize("tune", [dict(table_scale=2), dict(table_scale=8), dict(small_max=64, small_group=4),
                                  dict(unroll=4, grid_per_cu=2), dict(small_max=512, table_scale=2),
                                  dict(table_scale=16), dict(waves=2), dict(waves=8),
                                  dict(waves=8, grid_per_cu=32), dict(waves=2, table_scale=2, small_max=512)])
def test_tuning_knobs_do_not_change_counts(lstc, oracle, tune):
    gs, want = _mixed(lstc, oracle)
    t = lstc.Tune(**tune)
    assert lstc.count_batch(gs, tune=t) == want
    assert lstc.count(gs[0], tune=t) == want[0]


def test_per_row_in_batch(lstc, oracle):
    gs, _ = _mixed(lstc, oracle)
    for g in gs[:3]:
        off, col = g.to_numpy()
        t_ref, rows_ref = oracle.count_k1(oracle.CSR(g.n, off, col))
        t, rows = lstc.count(g, per_row=True)
        assert t == t_ref and np.array_equal(rows, rows_ref)
