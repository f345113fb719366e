"""GPU tests of the serving loop (lstc_server_*): exact results alone and in the
mix, HBM-budget admission and rejection, caller-resident CSR requests."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _req_list(rng, k, lo=1e3, hi=2e5):
    out = []
    for i in range(k):
        m = int(round(10 ** rng.uniform(np.log10(lo), np.log10(hi))))
        if rng.uniform() < 0.7:
            scale = max(6, int(np.ceil(np.log2(max(m / 8, 64)))))
            out.append(("rmat", scale, m, 1000 + i))
        else:
            n = max(64, m // 10)
            out.append(("er", n, m, 2000 + i))
    return out


def _oracle_count(O, kind, a, m, seed):
    if kind == "rmat":
        n, s, d = O.gen_rmat(a, seed=seed, m_raw=m)
    else:
        n, s, d = O.gen_er(a, seed=seed, m_raw=m)
    sym, g = O.clean_orient(n, s, d)
    return O.count(g), g.m


def _submit(srv, req):
    kind, a, m, seed = req
    return srv.submit_rmat(a, m, seed) if kind == "rmat" else srv.submit_er(a, m, seed)


def test_results_exact_alone_and_in_mix(lstc, oracle):
    rng = np.random.default_rng(5)
    reqs = _req_list(rng, 40)
    want = [_oracle_count(oracle, *r) for r in reqs]
    srv = lstc.Server(2 << 30)
    try:
        # in the mix: staggered arrivals so later requests join while earlier ones run
        ids = []
        for r in reqs:
            ids.append(_submit(srv, r))
            time.sleep(0.0005)
        got = [srv.wait(i, 120) for i in ids]
        assert [g["count"] for g in got] == [w[0] for w in want]
        assert [g["edges"] for g in got] == [w[1] for w in want]
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["rejected"] == 0
        assert st["pool_used_bytes"] == 0
        # alone: one request at a time through the same server
        for r, w in zip(reqs[:8], want[:8]):
            assert srv.wait(_submit(srv, r), 60)["count"] == w[0]
    finally:
        srv.close()


def test_over_budget_rejected(lstc, oracle):
    srv = lstc.Server(8 << 20)  # 8 MiB
    try:
        with pytest.raises(lstc.LstcError) as e:
            srv.submit_rmat(20, 16 << 20, 1)
        assert e.value.status == -3  # LSTC_ERR_OVER_BUDGET
        assert srv.stats()["rejected"] == 1
        with pytest.raises(lstc.LstcError) as e:  # invalid quadrant probabilities: refused, the server runs on
            srv.submit_rmat(10, 1000, 1, a=0.6, b=0.3, c=0.3)
        assert e.value.status == -1  # LSTC_ERR_INVALID
        rid = srv.submit_er(500, 3000, 7)  # a small one still runs
        assert srv.wait(rid, 60)["count"] == _oracle_count(oracle, "er", 500, 3000, 7)[0]
    finally:
        srv.close()


def test_tight_budget_queues_and_stays_exact(lstc, oracle):
    rng = np.random.default_rng(11)
    reqs = _req_list(rng, 24, 2e4, 1e5)
    want = [_oracle_count(oracle, *r)[0] for r in reqs]
    srv = lstc.Server(48 << 20)  # room for only a few working sets at a time
    try:
        ids = [_submit(srv, r) for r in reqs]
        got = [srv.wait(i, 120)["count"] for i in ids]
        assert got == want
        st = srv.stats()
        assert st["batches"] >= 2 and st["pool_used_bytes"] == 0
    finally:
        srv.close()


def test_resident_csr_requests(lstc, oracle):
    graphs, want = [], []
    for seed in range(5):
        n, s, d = oracle.gen_rmat(11, 8, seed=seed)
        sym, g = oracle.clean_orient(n, s, d)
        graphs.append(lstc.Graph.from_numpy(g.off, g.col))
        want.append(oracle.count(g))
    srv = lstc.Server(256 << 20)
    try:
        ids = [srv.submit_graph(g) for g in graphs]
        assert [srv.wait(i, 60)["count"] for i in ids] == want
        assert [lstc.count(g) for g in graphs] == want
    finally:
        srv.close()


def test_poll_not_ready_then_ready(lstc, oracle):
    srv = lstc.Server(512 << 20)
    try:
        rid = srv.submit_rmat(14, 1 << 18, 3)
        seen = None
        t0 = time.time()
        while seen is None and time.time() - t0 < 60:
            seen = srv.poll(rid)
        assert seen is not None
        assert seen["count"] == _oracle_count(oracle, "rmat", 14, 1 << 18, 3)[0]
        assert srv.poll(rid) is None  # a taken result is gone
    finally:
        srv.close()


def test_cancel_queued_request(lstc):
    """A request not admitted yet can be cancelled (its result is LSTC_ERR_CANCELLED, it takes no
    memory); admitted or finished ones keep their exact result; unknown ids are refused."""
    import json
    import os
    from trianglecount_amd._lib import LSTC_ERR_CANCELLED
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "rmat_counts.json")))["18"]
    # ~290 MB per RMAT s18 working set: the pool (7/8 of 480 MB) holds one at a time,
    # so the second and third requests wait in the queue while the first is counted
    srv = lstc.Server(480 << 20)
    try:
        ids = [srv.submit_rmat(18, 16 << 18, gold["seed"]) for _ in range(3)]
        assert srv.cancel(ids[2]) is True
        for rid in ids[:2]:
            r = srv.wait(rid, 120.0)
            assert r["count"] == gold["triangles"] and r["edges"] == gold["m"]
        with pytest.raises(lstc.LstcError) as e:
            srv.wait(ids[2], 120.0)
        assert e.value.status == LSTC_ERR_CANCELLED
        assert srv.cancel(ids[0]) is False  # finished: nothing to cancel
        with pytest.raises(lstc.LstcError):
            srv.cancel(ids[2] + 1000)  # never returned by this server
        st = srv.stats()
        assert st["completed"] == 3
    finally:
        srv.close()


def test_grouped_preparation_exact(lstc, oracle):
    """A burst of generated requests is prepared in groups (their union through one
    clean/orient pipeline, then split back into one CSR per graph with local ids):
    every count and edge total equals the oracle's, with caller CSRs interleaved
    (a group stops at one) and a budget that forces some groups to be halved."""
    rng = np.random.default_rng(21)
    reqs = _req_list(rng, 150, 1e3, 5e4)
    want = [_oracle_count(oracle, *r) for r in reqs]
    _, og = oracle.clean_orient(*oracle.gen_rmat(10, 8, seed=4))
    g = lstc.Graph.from_numpy(og.off, og.col)
    csr_want = (oracle.count(og), og.m)
    for budget in (4 << 30, 96 << 20):
        srv = lstc.Server(budget)
        try:
            ids, exp = [], []
            for i, r in enumerate(reqs):
                ids.append(_submit(srv, r))
                exp.append(want[i])
                if i in (40, 100):
                    ids.append(srv.submit_graph(g))
                    exp.append(csr_want)
            got = [srv.wait(x, 120) for x in ids]
            assert [(r["count"], r["edges"]) for r in got] == exp, budget
            st = srv.stats()
            assert st["completed"] == len(ids) and st["pool_used_bytes"] == 0
            assert st["max_prep_graphs"] > 1 and st["prep_groups"] < len(reqs), st
        finally:
            srv.close()


def test_grouped_preparation_degenerate_requests(lstc, oracle):
    """Requests with no raw edges, only loops or a single edge, prepared in the same groups as
    ordinary ones (a burst): every count and edge total equals the oracle's."""
    odd = [("rmat", 1, 0, 5), ("er", 1, 5, 6), ("rmat", 3, 1, 7), ("er", 2, 0, 8), ("rmat", 2, 3, 9)]
    rng = np.random.default_rng(31)
    reqs = []
    for i, r in enumerate(_req_list(rng, 40, 1e3, 2e4)):
        reqs.append(r)
        reqs.append(odd[i % len(odd)])
    want = [_oracle_count(oracle, *r) for r in reqs]
    srv = lstc.Server(1 << 30)
    try:
        ids = [_submit(srv, r) for r in reqs]
        got = [srv.wait(x, 120) for x in ids]
        assert [(r["count"], r["edges"]) for r in got] == want
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["pool_used_bytes"] == 0 and st["max_prep_graphs"] > 1, st
    finally:
        srv.close()
