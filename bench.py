"""Benchmark: exact-count GTEPS at RMAT s23 ef16 on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--scale 23]

A step is one exact count (lstc_count_split_stream: plan + count kernel) of the
resident oriented CSR of a synthetic RMAT graph (scale 23, edge factor 16,
a/b/c/d .57/.19/.19/.05, seed 1), generated, cleaned and oriented on the GPU
before timing, with the in-lists, segment words and row-aligned streaming copy
of the columns its preprocessing emits.
GTEPS = undirected edges after cleaning / seconds per count / 1e9.

Besides the headline the line carries the partitioned workload (8 halo parts
in one batched launch), the mixed-shape batch (batched vs per-graph), the
serving stream (p50/p99, every count checked) and a serving load sweep.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
GOLD = os.path.join(ROOT, "tests", "golden")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def graph_stats(g):
    import torch
    off = g.off[: g.n + 1]
    d = (off[1:] - off[:-1]).to(torch.int64)
    indeg = torch.bincount(g.col[: g.m].to(torch.int64), minlength=g.n)[: g.n]
    return {"n": g.n, "m": g.m, "max_out": int(d.max().item()) if g.n else 0,
            "sum_out_sq": int((d * d).sum().item()), "sum_in_out": int((indeg * d).sum().item())}


def merge_model_bytes(st):
    """Merge model of the reference's kernel: 8(n+1) + 4m + sum_{(u,v) in E+} [16 + 4(d+(u) + d+(v))]."""
    return 8 * (st["n"] + 1) + 4 * st["m"] + 16 * st["m"] + 4 * (st["sum_out_sq"] + st["sum_in_out"])


def committed_traffic(scale):
    """HBM-side bytes per count-kernel launch from the committed rocprofv3 PMC summary of this
    workload (profiles/k_count_rmat_s<scale>_pmc.txt: FETCH_SIZE and WRITE_SIZE per dispatch,
    each scaled by what the counter reads for a known-size copy kernel in the same run), or None."""
    p = os.path.join(ROOT, "profiles", f"k_count_rmat_s{scale}_pmc.txt")
    try:
        with open(p) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":", 1)[1]), os.path.relpath(p, ROOT)
    except OSError:
        pass
    return None, None


def timed_counts(T, g, steps, warmup, stream, ws, split=True):
    """Step time (plan + count kernel) from hipEvents on the launch stream, and the count
    kernel's own time from the events the library records around each launch."""
    import torch
    out = torch.zeros(1, dtype=torch.int64, device="cuda")
    for _ in range(warmup):
        T.count(g, out=out, ws=ws, sync=False, split=split, stream=stream)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    with T.KernelTimer() as kt:
        e0.record(stream)
        for _ in range(steps):
            T.count(g, out=out, ws=ws, sync=False, split=split, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / steps
    assert kt.launches == steps, (kt.launches, steps)
    return int(out[0].item()), ms, kt.ms / steps, wall


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(g, run_s=2.0, runs=5):
    """Oracle merge counter (OpenMP) on strided row samples of the same oriented CSR.

    All cores: every k-th row, k calibrated so one run takes about run_s seconds,
    median of `runs` runs (OMP_PROC_BIND=close).  One thread: a 16x sparser
    sample the same way, median of 3 runs.  Rates are sampled oriented edges / s.
    """
    import statistics
    import numpy as np
    os.environ.setdefault("OMP_PROC_BIND", "close")  # before the OpenMP runtime starts
    from oracle import oracle as O
    off, col = g.to_numpy()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    n = g.n

    def rate(nthreads, target_s, reps, sparsity):
        stride = max(1, n // 2000) * sparsity
        O.count_rows(off, col, np.arange(0, n, stride * 8, dtype=np.uint32), nthreads)  # thread pool start-up
        t0 = time.perf_counter()
        O.count_rows(off, col, np.arange(0, n, stride, dtype=np.uint32), nthreads)
        dt = time.perf_counter() - t0
        stride = max(1, int(round(stride * max(dt, 1e-4) / target_s)))
        rows = np.arange(0, n, stride, dtype=np.uint32)
        edges = int((off[rows.astype(np.int64) + 1] - off[rows.astype(np.int64)]).sum())
        tri = O.count_rows(off, col, rows, nthreads)  # warm-up
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            O.count_rows(off, col, rows, nthreads)
            times.append(time.perf_counter() - t0)
        med = statistics.median(times)
        return edges / med / 1e9, stride, rows, edges, med, tri

    v_all, st_all, rows_all, e_all, t_all, tri_all = rate(threads, run_s, runs, 1)
    v_one, st_one, rows_one, e_one, t_one, _ = rate(1, run_s, 3, 16)
    return {"value": round(v_all, 5), "unit": "GTEPS", "cores": threads, "kind": "port",
            "sample": f"every {st_all}th of {n} rows of the same oriented CSR ({len(rows_all)} rows, {e_all} oriented "
                      f"edges) counted by the oracle merge counter, median of {runs} runs {t_all:.2f} s",
            "single_thread_value": round(v_one, 6),
            "single_thread_sample": f"every {st_one}th row ({len(rows_one)} rows, {e_one} oriented edges), median of 3 "
                                    f"runs {t_one:.2f} s",
            "host": {"nproc": os.cpu_count(), "cpu_model": _cpu_model(),
                     "omp_proc_bind": os.environ.get("OMP_PROC_BIND")}}, (rows_all, tri_all)


def sample_matches_gpu(T, g, rows, tri):
    """The CPU sample's count equals the sum of the GPU's per-row partials g_sum[u] over the same rows."""
    import numpy as np
    _, per_row = T.count(g, per_row=True, split=False)
    return int(per_row[rows.astype(np.int64)].sum(dtype=np.uint64)) == tri


def autotuned(T, g, ws, stream, steps=5):
    """lstc_autotune on this graph: the tune it picks and the count's time with it vs the defaults
    (both timed the same way, plan + count kernel per step)."""
    tune, _, _ = T.autotune([g], reps=3)
    out = {"tune": tune.as_dict()}
    _, ms0, k0, _ = timed_counts(T, g, steps, 1, stream, ws)
    import torch
    o = torch.zeros(1, dtype=torch.int64, device="cuda")
    for _ in range(2):
        T.count(g, out=o, ws=ws, sync=False, tune=tune, stream=stream)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        T.count(g, out=o, ws=ws, sync=False, tune=tune, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    out.update(default_ms=round(ms0, 4), tuned_ms=round(e0.elapsed_time(e1) / steps, 4))
    return out


def _replay_ms(T, graphs, steps):
    """Device time of one persistent batched launch of ``graphs`` (hipGraph replay); returns (ms, counts)."""
    import torch
    plan = T.BatchPlan(graphs)
    plan.launch_graph()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        plan.launch_graph()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    per = plan.results()
    plan.close()
    return ms, per


def _gold(name):
    try:
        return json.load(open(os.path.join(GOLD, name)))
    except OSError:
        return None


def partitioned(T, g, k, seed, steps, ws):
    """RMAT halo parts (k-way seeded random assignment) counted in one persistent launch.

    Each part carries owned-source in-lists (split counting); the same parts
    without them (plain CSR counting) are timed beside.  Part counts are checked
    against the oracle's committed counts when they exist for this workload.
    """
    import torch
    part = T.partition_random(g.n, k, seed=seed)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    parts = T.halo_extract_parts(g, part, k, mode=0, ws=ws)
    torch.cuda.synchronize()
    t_halo = time.perf_counter() - t0
    t0 = time.perf_counter()
    for h in parts:
        T.in_lists(h, ws=ws)
    torch.cuda.synchronize()
    t_inl = time.perf_counter() - t0
    ms, per = _replay_ms(T, parts, steps)
    for h in parts:
        h.in_off = h.in_col = h.in_owned = h.col_seg = h.in_seg = None
    ms_plain, per_plain = _replay_ms(T, parts, max(1, steps // 2))
    if per_plain != per:
        raise SystemExit(f"split part counts {per} differ from plain {per_plain}")
    gold = _gold(f"halo_s{g.meta.get('scale')}_k{k}.json")
    checked = None
    if gold and gold.get("seed") == g.meta.get("seed") and gold.get("partition_seed") == seed:
        checked = per == [q["triangles"] for q in gold["parts"]]
        if not checked:
            raise SystemExit(f"part counts {per} differ from the oracle's {[q['triangles'] for q in gold['parts']]}")
    return {"k": k, "mode": "P u N(P) (reference un-pruned halo)", "ms": round(ms, 3),
            "gteps": round(g.m / (ms / 1e3) / 1e9, 4), "sum_parts": sum(per),
            "per_part_oracle_checked": checked,
            "plain_csr_ms": round(ms_plain, 3), "halo_extract_s": round(t_halo, 4),
            "in_lists_s": round(t_inl, 4), "part_rows": [p.n for p in parts],
            "part_owned": [p.owned for p in parts], "part_arcs": [p.m for p in parts]}


def support_leg(T, g, tri, st, work, steps=3):
    """Per-edge support and per-vertex triangle counts of the headline graph: the split rule
    (lstc_triangle_support_split, the graph's in-lists) and the plain walk (lstc_triangle_support);
    time per call, and the sums every triangle fixes (it adds one to three arcs and three vertices).
    Algorithmic bytes: the streamed list entries (4 B each: the count's streamed entries for the
    split rule, sum over vertices of d-(v) d+(v) for the plain walk), the tables' keys, two row
    offsets per located list, the outputs written once, and one 8-byte read-modify-write per
    triangle for the support of its streamed arc."""
    import torch
    out = {}
    for name, split in (("split", True), ("plain", False)):
        T.triangle_support(g, split=split)  # first call outside the timing (output allocations)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            sup, pv = T.triangle_support(g, split=split)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        s_sup = int(sup.to(torch.int64).sum().item())
        s_pv = int(pv.sum().item())
        if split:
            alg = 4 * (work["streamed"] + work["hashed_keys"]) + 16 * (work["segments"] + work["tiny_segments"])
        else:
            alg = 4 * st["sum_in_out"] + 4 * g.m + 16 * g.m
        alg += 4 * g.m + 8 * g.n + 8 * tri
        out[name] = {"ms": round(ms, 3), "gteps": round(g.m / (ms / 1e3) / 1e9, 4), "steps": steps,
                     "algorithmic_bytes": alg, "achieved_gbs": round(alg / (ms / 1e3) / 1e9, 1),
                     "frac": round(alg / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                     "support_sum": s_sup, "per_vertex_sum": s_pv,
                     "sums_equal_3x_triangles": s_sup == 3 * tri == s_pv}
        del sup, pv
    return out


def _per_graph_calls(T, graphs, ws, tune=None):
    """Pre-bound C calls of the per-graph path: graph i counts into res[i] (lstc_count_split
    when it carries in-lists, else lstc_count), sharing one workspace sized for the largest."""
    import ctypes as C
    import torch
    L = T.load()
    need = 0
    tp = C.byref(tune) if tune is not None else None
    for g in graphs:
        nb = C.c_size_t()
        T.check(L.lstc_count_workspace(g.n, g.owned, g.m, tp, C.byref(nb)), "lstc_count_workspace")
        need = max(need, nb.value)
    wp, wb = ws.get(need)
    res = torch.zeros(len(graphs), dtype=torch.int64, device="cuda")
    sh = C.c_void_p(T.stream_handle(None))
    calls = []
    for i, g in enumerate(graphs):
        out = C.c_void_p(res.data_ptr() + 8 * i)
        if g.split_ready() and g.scol is not None and g.col_seg is not None:  # words point into the copy
            calls.append((L.lstc_count_split_stream,
                          (C.c_void_p(g.off.data_ptr()), C.c_void_p(g.col.data_ptr()),
                           C.c_void_p(g.in_off.data_ptr()), C.c_void_p(g.in_col.data_ptr()),
                           C.c_void_p(g.col_seg.data_ptr()), C.c_void_p(g.in_seg.data_ptr()),
                           C.c_void_p(g.scol.data_ptr()), int(g.scol.numel()), g.n, g.owned, g.m, out, tp, wp, wb,
                           sh)))
        elif g.split_ready():
            seg = [C.c_void_p(t.data_ptr()) if t is not None else None for t in (g.col_seg, g.in_seg)]
            calls.append((L.lstc_count_split, (C.c_void_p(g.off.data_ptr()), C.c_void_p(g.col.data_ptr()),
                                               C.c_void_p(g.in_off.data_ptr()), C.c_void_p(g.in_col.data_ptr()),
                                               seg[0], seg[1], g.n, g.owned, g.m, out, tp, wp, wb, sh)))
        else:
            calls.append((L.lstc_count, (C.c_void_p(g.off.data_ptr()), C.c_void_p(g.col.data_ptr()), g.n, g.owned,
                                         g.m, out, None, tp, wp, wb, sh)))
    return calls, res


def _per_graph(T, graphs, ws, reps, tune=None):
    """Wall time of the per-graph loop (one launch, then a synchronising copy of its count, per
    graph) and its summed device time; returns (wall_ms, device_ms, counts)."""
    import torch
    calls, res = _per_graph_calls(T, graphs, ws, tune)
    for fn, args in calls[:3]:
        T.check(fn(*args), "count")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        counts = []
        for i, (fn, args) in enumerate(calls):
            T.check(fn(*args), "count")
            counts.append(int(res[i].item()))  # synchronising 8-byte copy, as cudaMemcpy in the loop
    wall = (time.perf_counter() - t0) / reps
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in calls]
    for (fn, args), (e0, e1) in zip(calls, ev):
        e0.record()
        T.check(fn(*args), "count")
        e1.record()
    torch.cuda.synchronize()
    return wall * 1e3, sum(e0.elapsed_time(e1) for e0, e1 in ev), counts


def batched_vs_per_graph(T, graphs, reps, ws):
    """Whole mixed workload, wall clock: the per-graph path (the reference's loop: per graph one
    count launch, then a synchronising copy of its count to the host; calls pre-bound so no
    Python-side allocation is timed) vs one persistent batched launch replayed from a hipGraph,
    counts copied back.  Both with the tune the autotuner picks for the batch, and the per-graph
    path also with the defaults.  Counts are checked against the oracle's committed counts."""
    import torch
    tune, tune_ms, _cached = T.autotune(graphs, reps=3)
    pg_ms, pg_dev_ms, per_graph = _per_graph(T, graphs, ws, reps, tune=tune)
    pg0_ms, pg0_dev_ms, per_graph0 = _per_graph(T, graphs, ws, reps, tune=None)
    plan = T.BatchPlan(graphs, tune=tune)
    plan.launch_graph()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        plan.launch_graph()
        batched = plan.results()
    t_batch = (time.perf_counter() - t0) / reps
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    plan.launch_graph()
    e1.record()
    torch.cuda.synchronize()
    batch_dev_ms = e0.elapsed_time(e1)
    plan.close()
    gold = _gold("mixed_counts.json")
    checked = None
    if gold:
        checked = batched == [q["triangles"] for q in gold["graphs"]]
    return {"graphs": len(graphs), "edges": int(sum(g.m for g in graphs)),
            "tune": tune.as_dict(),
            "per_graph_ms": round(pg_ms, 3), "batched_ms": round(t_batch * 1e3, 3),
            "speedup": round(pg_ms / (t_batch * 1e3), 3),
            "per_graph_device_ms": round(pg_dev_ms, 3), "batched_device_ms": round(batch_dev_ms, 3),
            "device_speedup": round(pg_dev_ms / batch_dev_ms, 3),
            "per_graph_default_tune_ms": round(pg0_ms, 3),
            "speedup_vs_default_tune_per_graph": round(pg0_ms / (t_batch * 1e3), 3),
            "identical": per_graph == batched == per_graph0, "oracle_checked": checked,
            "triangles": int(sum(batched))}


def _serve(T, reqs, budget_gb, t_scale=1.0):
    """Submits ``reqs`` at their arrival times (scaled by t_scale; 0 = all at once) and waits for all."""
    import numpy as np
    srv = T.Server(int(budget_gb * (1 << 30)))
    ids = []
    with T.KernelTimer() as kt:  # the server's count kernels, on its count stream
        t0 = time.perf_counter()
        for r in reqs:
            at = r.arrive_s * t_scale
            while time.perf_counter() - t0 < at:
                pass
            ids.append(srv.submit_rmat(r.a, r.m_raw, r.seed) if r.kind == "rmat" else
                       srv.submit_er(r.a, r.m_raw, r.seed))
        res = [srv.wait(i) for i in ids]
        wall = time.perf_counter() - t0
        st = srv.stats()
        srv.close()
    lat = np.array([x["latency_s"] for x in res])
    edges = int(sum(x["edges"] for x in res))
    return res, {"wall_s": round(wall, 3), "edges": edges, "edges_per_s": round(edges / wall, 1),
                 "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                 "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
                 "mean_ms": round(float(lat.mean()) * 1e3, 3), "max_ms": round(float(lat.max()) * 1e3, 3),
                 "batches": int(st["batches"]), "graphs_per_launch": round(len(reqs) / max(1, st["batches"]), 2),
                 "count_kernel_busy": round(kt.ms / 1e3 / wall, 3),
                 "max_batch_graphs": int(st["max_batch_graphs"]), "rejected": int(st["rejected"]),
                 "prep_groups": int(st["prep_groups"]),
                 "graphs_per_prep": round(len(reqs) / max(1, st["prep_groups"]), 2),
                 "max_prep_graphs": int(st["max_prep_graphs"])}


def serving(T, n_req, rate, budget_gb, seed, verify):
    """Serving loop: the seeded Poisson stream of count requests through lstc_server.

    Request i: raw edge count E_i log-uniform in [1e3, 5e6]; 70% RMAT, 30% ER
    (trianglecount_amd.workloads.serving_stream).  Latency runs from submission
    to the count being available on the host.

    verify: "all" checks every count and edge total, against the oracle's committed
    counts (tests/golden/serving_counts.json) when they cover this stream, else by
    running the CPU oracle on every request; "none" skips the check.
    """
    from trianglecount_amd.workloads import serving_stream
    reqs = serving_stream(n_req, rate, seed)
    res, out = _serve(T, reqs, budget_gb)
    out = {"requests": n_req, "rate_per_s": rate, "offered_edges_per_s": None, "hbm_budget_gb": budget_gb,
           "mix": "70% RMAT / 30% ER, raw edges log-uniform [1e3, 5e6]",
           "raw_edges": int(sum(r.m_raw for r in reqs)), **out}
    out["offered_edges_per_s"] = round(out["edges"] / reqs[-1].arrive_s, 1)
    out["gteps"] = round(out["edges_per_s"] / 1e9, 4)
    if verify == "none":
        return out
    gold = _gold("serving_counts.json")
    if gold and (gold["requests"], gold["rate_per_s"], gold["seed"]) == (n_req, rate, seed):
        bad = sum((r["count"], r["edges"]) != (gold["triangles"][i], gold["edges"][i]) for i, r in enumerate(res))
        out["verified"] = n_req
        out["verified_against"] = "tests/golden/serving_counts.json (oracle counts of every request)"
        out["mismatches"] = int(bad)
        return out
    import concurrent.futures as cf
    from oracle import oracle as O

    def check(i):
        r = reqs[i]
        n, src, dst = O.gen_rmat(r.a, 16, seed=r.seed, m_raw=r.m_raw) if r.kind == "rmat" else \
            O.gen_er(r.a, 20.0, seed=r.seed, m_raw=r.m_raw)
        _, g = O.clean_orient(n, src, dst)
        return O.count(g, threads=1) == res[i]["count"] and g.m == res[i]["edges"]

    t1 = time.perf_counter()
    with cf.ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
        ok = list(ex.map(check, range(n_req)))
    out["verified"] = n_req
    out["verified_against"] = "CPU oracle run on every request"
    out["mismatches"] = int(len(ok) - sum(ok))
    out["verify_s"] = round(time.perf_counter() - t1, 1)
    return out


def serving_load(T, n_req, rates, budget_gb, seed):
    """Load sweep: the first n_req requests of the same stream at increasing arrival rates, then
    all at once (closed burst); each point's throughput, latency and graphs per launch."""
    from trianglecount_amd.workloads import serving_stream
    base = serving_stream(10000, 1000.0, seed)[:n_req]  # the head of the committed stream
    gold = _gold("serving_counts.json")
    points = []
    for rate in list(rates) + [None]:
        res, p = _serve(T, base, budget_gb, t_scale=0.0 if rate is None else 1000.0 / rate)
        p = {"rate_per_s": rate if rate is not None else "burst", **p}
        p["offered_edges_per_s"] = None if rate is None else round(p["edges"] / (base[-1].arrive_s * 1000.0 / rate), 1)
        if gold and gold["seed"] == seed:
            p["mismatches"] = int(sum(r["count"] != gold["triangles"][i] for i, r in enumerate(res)))
        points.append(p)
        log(f"serving load {p}")
    return {"requests_per_point": n_req, "points": points,
            "capacity_edges_per_s": max(p["edges_per_s"] for p in points)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scale", type=int, default=23)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the partitioned, mixed and serving workloads")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the timed split count (for profiling: no plain, extra or CPU legs)")
    ap.add_argument("--serve-requests", type=int, default=10000)
    ap.add_argument("--serve-rate", type=float, default=1000.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--serve-budget-gb", type=float, default=64.0)
    ap.add_argument("--serve-verify", choices=["all", "none"], default="all")
    ap.add_argument("--load-requests", type=int, default=2000, help="requests per point of the serving load sweep")
    ap.add_argument("--load-rates", default="2000,4000,8000", help="arrival rates of the load sweep (+ a burst)")
    ap.add_argument("--legs", default="autotune,partitioned,support,mixed,serving,load",
                    help="the extra workloads to run (comma list; for profiling one of them)")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="first run an int32 elementwise copy of this many MiB (reads N, writes N bytes), "
                         "the known-size kernel that calibrates HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(a sharded count would need only one 8-byte sum per graph across GPUs)", file=sys.stderr)
        sys.exit(2)
    import torch
    import trianglecount_amd as T
    from trianglecount_amd import workloads as W
    T.load()
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    ws = T.Workspace()
    if args.headline_only:
        args.no_extra = args.no_cpu = True
    if args.copy_calib_mib:
        a = torch.ones(args.copy_calib_mib << 18, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        torch.add(a, 1, out=b)
        torch.cuda.synchronize()
        log(f"copy calibration: {args.copy_calib_mib} MiB read + {args.copy_calib_mib} MiB written")
        del a, b

    t0 = time.perf_counter()
    n, s, d = T.generate_rmat(args.scale, 16, seed=args.seed)
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t0
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    # a first run of the same size outside the timing: workspace and output buffers come from the
    # allocator's cache in the timed run, as they do for a caller that prepares graphs repeatedly
    T.clean_orient(n, s, d, ws=ws, in_lists=True, stream_copy=True)
    torch.cuda.synchronize()
    e0.record()
    # the count's input: oriented CSR, in-lists, segment words and the streaming copy they point into
    g = T.clean_orient(n, s, d, ws=ws, in_lists=True, stream_copy=True)
    e1.record()
    torch.cuda.synchronize()
    t_prep = e0.elapsed_time(e1) / 1e3
    g.meta.update(scale=args.scale, seed=args.seed)
    t_sym = None
    if not args.no_extra and "partitioned" in args.legs:  # the halo extraction reads the symmetric CSR
        e0.record()
        gs = T.clean_orient(n, s, d, ws=ws, keep_symmetric=True)
        e1.record()
        torch.cuda.synchronize()
        t_sym = e0.elapsed_time(e1) / 1e3
        g.sym_off, g.sym_col = gs.sym_off, gs.sym_col
        del gs
    del s, d
    st = graph_stats(g)
    log(f"rmat s{args.scale}: n={g.n} m={g.m} max d+={st['max_out']} gen {t_gen:.2f}s prep {t_prep*1e3:.1f} ms")

    tri, ms, kern_ms, wall = timed_counts(T, g, args.steps, args.warmup, stream, ws)
    work = T.work_stats([g], ws=ws)[0]
    tri_plain, ms_plain = tri, None
    inl_ms = None
    if not args.headline_only:
        tri_plain, ms_plain, _, _ = timed_counts(T, g, max(1, args.steps // 2), 1, stream, ws, split=False)
        log(f"plain-CSR count (no in-lists) {ms_plain:.3f} ms/step")
        # drop-in cost: in-lists and segment words built from a bare oriented CSR
        bare = T.Graph(g.n, g.m, g.off, g.col)
        ea, eb = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        T.in_lists(bare, ws=ws)
        torch.cuda.synchronize()
        ea.record()
        T.in_lists(bare, ws=ws)
        eb.record()
        torch.cuda.synchronize()
        inl_ms = ea.elapsed_time(eb)
        del bare
    if tri_plain != tri:
        raise SystemExit(f"split count {tri} differs from the plain-CSR count {tri_plain}")
    gold = (_gold("rmat_counts.json") or {}).get(str(args.scale))
    if gold and args.seed == gold["seed"] and tri != gold["triangles"]:
        raise SystemExit(f"count {tri} differs from the oracle's {gold['triangles']}")
    gteps = g.m / (ms / 1e3) / 1e9
    achieved = work["bytes"] / (kern_ms / 1e3) / 1e9
    mm = merge_model_bytes(st)
    traffic, traffic_src = committed_traffic(args.scale)
    log(f"count {tri} in {ms:.3f} ms/step (kernel {kern_ms:.3f} ms) -> {gteps:.3f} GTEPS; "
        f"algorithmic {work['bytes']/1e9:.1f} GB -> {achieved:.0f} GB/s")

    extra = {}
    legs = set() if args.no_extra else {x for x in args.legs.split(",") if x}
    if "autotune" in legs:
        tuned = {f"s{args.scale}": autotuned(T, g, ws, stream)}
        n21, s21, d21 = T.generate_rmat(21, 16, seed=args.seed)
        g21 = T.clean_orient(n21, s21, d21, ws=ws, in_lists=True)
        del s21, d21
        tuned["s21"] = autotuned(T, g21, ws, stream)
        del g21
        extra["autotune"] = tuned
        log(f"autotune: {tuned}")
    if "partitioned" in legs:
        extra["partitioned"] = partitioned(T, g, 8, args.seed, max(1, args.steps // 2), ws)
        log(f"partitioned k=8: {extra['partitioned']}")
        if extra["partitioned"]["sum_parts"] != tri:
            raise SystemExit("partitioned sum differs from the unpartitioned count")
    g.sym_off = g.sym_col = None
    torch.cuda.empty_cache()
    if "support" in legs:
        extra["support"] = support_leg(T, g, tri, st, work)
        log(f"triangle support: {extra['support']}")
        if not all(v["sums_equal_3x_triangles"] for v in extra["support"].values()):
            raise SystemExit("support / per-vertex sums differ from 3x the count")
        torch.cuda.empty_cache()
    if "mixed" in legs:
        mg, _names = W.mixed_graphs(args.seed, ws)
        torch.cuda.synchronize()
        extra["mixed"] = batched_vs_per_graph(T, mg, 3, ws)
        log(f"mixed workload: {extra['mixed']}")
        if not extra["mixed"]["identical"] or extra["mixed"]["oracle_checked"] is False:
            raise SystemExit("batched counts differ from per-graph or oracle counts")
        del mg
        torch.cuda.empty_cache()
    if "serving" in legs:
        extra["serving"] = serving(T, args.serve_requests, args.serve_rate, args.serve_budget_gb, args.seed,
                                   args.serve_verify)
        log(f"serving: {extra['serving']}")
        if extra["serving"].get("mismatches"):
            raise SystemExit("served counts differ from the oracle")
    if "load" in legs:
        rates = [float(x) for x in args.load_rates.split(",") if x]
        extra["serving_load"] = serving_load(T, args.load_requests, rates, args.serve_budget_gb, args.seed)
        if any(p.get("mismatches") for p in extra["serving_load"]["points"]):
            raise SystemExit("served counts under load differ from the oracle")
    cpu = None
    if not args.no_cpu:
        cpu, (rows, tri_rows) = cpu_baseline(g)
        cpu["sample_matches_gpu_per_row"] = sample_matches_gpu(T, g, rows, tri_rows)
        log(f"cpu baseline {cpu['value']} GTEPS on {cpu['cores']} threads, {cpu['single_thread_value']} GTEPS on one "
            f"({cpu['host']['cpu_model']}, nproc {cpu['host']['nproc']}); sample == GPU per-row partials: "
            f"{cpu['sample_matches_gpu_per_row']}")
    res = {
        "metric": "exact-count GTEPS at RMAT s23 ef16; batched vs per-graph launch; p99 latency",
        "value": round(gteps, 4), "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32/u64 (exact integer)",
        "data": "synthetic (seeded RMAT generated on the GPU)",
        "config": {"workload": f"rmat-s{args.scale}-ef16 unpartitioned exact count", "scale": args.scale,
                   "edge_factor": 16, "abcd": [0.57, 0.19, 0.19, 0.05], "seed": args.seed,
                   "n": g.n, "m": g.m, "triangles": tri, "oracle_checked": bool(gold),
                   "prep_ms": round(t_prep * 1e3, 2),
                   "prep_detail": "clean + orient + in-lists + segment words + row-aligned streaming copy "
                                  "(lstc_clean_orient_stream), one synchronisation",
                   "symmetric_prep_ms": t_sym and round(t_sym * 1e3, 2),
                   "end_to_end_gteps": round(g.m / (t_prep + ms / 1e3) / 1e9, 4),
                   "count_mode": "split, lists streamed from the row-aligned copy (lstc_count_split_stream)",
                   "in_lists_from_bare_csr_ms": inl_ms and round(inl_ms, 3),
                   "plain_csr_count_ms": ms_plain and round(ms_plain, 4),
                   "plain_csr_gteps": ms_plain and round(g.m / (ms_plain / 1e3) / 1e9, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_count", "kernel_ms": round(kern_ms, 4),
                     "algorithmic_bytes": work["bytes"], "work": work,
                     "l2_fabric_gbs": traffic and round(traffic / (kern_ms / 1e3) / 1e9, 1),
                     "traffic_source": traffic_src,
                     "merge_model_bytes": mm, "merge_model_gbs": round(mm / (kern_ms / 1e3) / 1e9, 1)},
        "cpu_baseline": cpu,
    }
    names = {"partitioned": "partitioned_s23_k8", "mixed": "batched_vs_per_graph", "serving": "serving",
             "serving_load": "serving_load", "autotune": "autotune", "support": "triangle_support_s23"}
    for k, v in extra.items():
        res[names[k]] = v
    print(json.dumps(res))


if __name__ == "__main__":
    main()
