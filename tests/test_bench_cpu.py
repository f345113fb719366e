"""bench.py's CPU baseline leg on a small graph (CPU only): the fields the bench line reports."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def test_cpu_baseline_fields(oracle):
    import bench
    n, s, d = oracle.gen_rmat(12, 16, seed=3)
    _, g = oracle.clean_orient(n, s, d)

    class Host:
        def __init__(self, g):
            self.n = g.n
            self._g = g

        def to_numpy(self):
            return self._g.off, self._g.col

    r, (rows, tri) = bench.cpu_baseline(Host(g), run_s=0.02, runs=3)
    # the sampled rows' count is the sum of the per-row partials g_sum[u] over the same rows
    _, per_row = oracle.count(g, per_row=True)
    assert tri == int(per_row[rows.astype(np.int64)].sum()) and len(rows) > 0
    assert r["unit"] == "GTEPS" and r["kind"] == "port"
    assert r["value"] > 0 and r["single_thread_value"] > 0
    assert r["cores"] >= 1 and r["host"]["nproc"] >= 1 and r["host"]["cpu_model"]
    assert "median of 3 runs" in r["sample"]
