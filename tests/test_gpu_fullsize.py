"""Full-size parity: the device pipeline at RMAT scales 21-23 (ef16, seed 1) against
the oracle's counts committed in tests/golden/rmat_counts.json.

generate -> clean + orient (+ in-lists) -> count, split and plain; at s21 also
the 8-part halo decomposition counted in one batched launch.
"""
import json
import os

import pytest

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "rmat_counts.json")))


@pytest.mark.parametrize("scale", [21, 22, 23])
def test_rmat_pipeline_matches_oracle(lstc, scale):
    want = GOLD[str(scale)]
    n, s, d = lstc.generate_rmat(scale, 16, seed=1)
    g = lstc.clean_orient(n, s, d, in_lists=True, keep_symmetric=scale == 21)
    del s, d
    assert (g.n, g.m) == (want["n"], want["m"])
    assert lstc.count(g) == want["triangles"]
    assert lstc.count(g, split=False) == want["triangles"]
    if scale == 21:
        part = lstc.partition_random(g.n, 8, seed=1)
        parts = [lstc.halo_extract(g, part, p, split=True) for p in range(8)]
        assert sum(lstc.count_batch(parts)) == want["triangles"]
