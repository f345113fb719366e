"""GPU tests of the autotuner (lstc_autotune) and its per-device cache."""
import pytest

import graphs as G

pytestmark = pytest.mark.gpu


def _batch(lstc, oracle):
    gs, want = [], []
    for n, s, d in [oracle.gen_rmat(13, 16, seed=2), oracle.gen_er(20000, 20, seed=3), G.complete(200),
                    oracle.gen_grid(60, 60, 0.3, 1)]:
        _, g = oracle.clean_orient(n, s, d)
        gs.append(lstc.Graph.from_numpy(g.off, g.col))
        want.append(oracle.count(g))
    return gs, want


def test_autotune_keeps_counts_and_caches(lstc, oracle, tmp_path):
    gs, want = _batch(lstc, oracle)
    cache = str(tmp_path / "tune.tsv")
    t, ms, hit = lstc.autotune(gs, reps=2, cache=cache)
    assert not hit and ms > 0
    d = t.as_dict()
    assert d["table_scale"] in (2, 4, 6, 8, 16) and d["unroll"] in (4, 8)
    assert lstc.count_batch(gs, tune=t) == want
    # the batch is under 2^23 arcs, so a default small_max runs capped at 128; the returned
    # tune must name what was timed, so replaying it resolves to itself
    assert lstc.tune_resolve(gs, t).as_dict() == d
    assert lstc.tune_resolve(gs).as_dict()["small_max"] == 128
    t2, ms2, hit2 = lstc.autotune(gs, reps=2, cache=cache)
    assert hit2 and ms2 is None and t2.as_dict() == d
    # another shape class misses the cache
    assert lstc.shape_class(gs[:1]) != lstc.shape_class(gs)
