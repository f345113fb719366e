"""Full-size GPU parity of the three large workloads against the oracle's committed counts.

* RMAT scale 23, ef16, seed 1, split into 8 seeded random parts; the un-pruned halo
  subgraph of every part extracted on the device and all 8 counted in ONE persistent
  batched launch (and its hipGraph replay): sizes, CSR checksums and owned-row counts
  per part equal tests/golden/halo_s23_k8.json, and the parts sum to the whole
  (the reference's per-part loop, src/WithPartition/without_sorting_tc.cu:244-451).
* The mixed-shape batch (2000 x 2000 grid + p = 0.3 diagonals, Chung-Lu 1M / 2.1 /
  mean 50 / w_max 5e4, 64 Erdos-Renyi graphs): every graph's count batched, per
  graph, with the autotuned tune and on hipGraph replay equals
  tests/golden/mixed_counts.json; the grid's count is twice its diagonals.
* The serving stream of 10,000 Poisson requests (raw edges log-uniform in
  [1e3, 5e6], 70% RMAT / 30% ER) through lstc_server with a 64 GB budget: every
  count and edge total equals tests/golden/serving_counts.json.

The fixtures are written by tests/golden/make_config_counts.py from the oracle.
"""
import json
import os
import time

import pytest
import torch

from trianglecount_amd import workloads as W

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _load(name):
    return json.load(open(os.path.join(GOLD, name)))


def _checksums(g):
    """The fingerprints of tests/golden/make_config_counts.py:csr_checksums, on the device."""
    off = g.off[: g.n + 1]
    col = g.col[: g.m].to(torch.int64) & 0xFFFFFFFF
    w = (torch.arange(g.m, device=col.device, dtype=torch.int64) & 1023) + 1
    return {"off_sum": int(off.sum().item()), "col_sum": int(col.sum().item()),
            "col_wsum": int((w * col).sum().item())}


def test_halo_s23_k8_one_batched_launch(lstc):
    gold = _load("halo_s23_k8.json")
    print(f"[halo s23 k8] generating", flush=True)
    n, s, d = lstc.generate_rmat(gold["scale"], gold["edge_factor"], seed=gold["seed"])
    g = lstc.clean_orient(n, s, d, keep_symmetric=True)
    del s, d
    assert (g.n, g.m) == (gold["n"], gold["m"])
    part = lstc.partition_random(g.n, gold["k"], seed=gold["partition_seed"])
    parts = lstc.halo_extract_parts(g, part, gold["k"], mode=gold["mode"], split=True)
    g.sym_off = g.sym_col = None
    for h, want in zip(parts, gold["parts"]):
        assert (h.n, h.owned, h.m) == (want["n_local"], want["n_owned"], want["m_local"]), want["part"]
        ck = _checksums(h)
        assert ck == {k: want[k] for k in ck}, want["part"]
    print(f"[halo s23 k8] parts extracted; one batched launch", flush=True)
    got = lstc.count_batch(parts)
    assert got == [q["triangles"] for q in gold["parts"]]
    assert sum(got) == gold["triangles"]
    plan = lstc.BatchPlan(parts)
    try:
        for _ in range(2):
            plan.counts.fill_(0)
            plan.launch_graph()
            assert plan.results() == got
    finally:
        plan.close()


def test_mixed_workload_batched_per_graph_autotuned(lstc):
    gold = _load("mixed_counts.json")
    want = [q["triangles"] for q in gold["graphs"]]
    graphs, names = W.mixed_graphs(gold["seed"])
    assert names == [q["name"] for q in gold["graphs"]]
    assert [(g.n, g.m) for g in graphs] == [(q["n"], q["m"]) for q in gold["graphs"]]
    assert want[0] == 2 * gold["graphs"][0]["diagonals"]  # grid: every triangle is half of one cell
    print(f"[mixed] {len(graphs)} graphs resident; batched", flush=True)
    assert lstc.count_batch(graphs) == want
    assert [lstc.count(g) for g in graphs] == want
    tune, ms, _ = lstc.autotune(graphs, reps=2)
    print(f"[mixed] autotuned {tune.as_dict()} {ms:.3f} ms", flush=True)
    assert lstc.tune_resolve(graphs, tune).as_dict() == tune.as_dict()
    assert lstc.count_batch(graphs, tune=tune) == want
    assert [lstc.count(g, tune=tune) for g in graphs] == want
    plan = lstc.BatchPlan(graphs, tune=tune)
    try:
        for _ in range(2):
            plan.counts.fill_(0)
            plan.launch_graph()
            assert plan.results() == want
    finally:
        plan.close()


def test_serving_stream_10000_requests(lstc):
    gold = _load("serving_counts.json")
    reqs = W.serving_stream(gold["requests"], gold["rate_per_s"], gold["seed"])
    assert "".join("R" if r.kind == "rmat" else "E" for r in reqs) == gold["kind"]
    assert [r.a for r in reqs] == gold["a"] and [r.m_raw for r in reqs] == gold["m_raw"]
    srv = lstc.Server(64 << 30)
    try:
        ids = []
        t0 = time.perf_counter()
        for r in reqs:  # Poisson arrivals at the stream's rate
            while time.perf_counter() - t0 < r.arrive_s:
                pass
            ids.append(srv.submit_rmat(r.a, r.m_raw, r.seed) if r.kind == "rmat" else srv.submit_er(r.a, r.m_raw,
                                                                                               r.seed))
        print(f"[serving] {len(ids)} submitted in {time.perf_counter() - t0:.1f} s", flush=True)
        res = [srv.wait(i, 300) for i in ids]
        st = srv.stats()
    finally:
        srv.close()
    bad = [i for i, r in enumerate(res) if (r["count"], r["edges"]) != (gold["triangles"][i], gold["edges"][i])]
    assert not bad, f"{len(bad)} requests differ, first {bad[:5]}"
    assert st["completed"] == len(reqs) and st["rejected"] == 0 and st["pool_used_bytes"] == 0
    assert max(r.m_raw for r in reqs) > 4_900_000  # the largest requests are in the checked stream
