"""Seeded definitions of the synthetic workloads the tests and the benchmark share.

These are parameter lists only (numpy, no device work), so the GPU tests, the
benchmark and the scripts that write the committed golden counts all derive
the same graphs from the same seed:

    mixed_specs(seed)          the mixed-shape batch: a 2000 x 2000 grid with
                               p = 0.3 cell diagonals, a Chung-Lu graph (1M
                               vertices, gamma 2.1, mean degree 50, w_max 5e4)
                               and 64 Erdos-Renyi graphs with n log-uniform in
                               [1e3, 1e5] and mean degree 20;
    serving_stream(n, rate, seed)
                               the serving stream: raw edge counts log-uniform
                               in [1e3, 5e6], 70% RMAT (a/b/c/d .57/.19/.19/.05)
                               and 30% Erdos-Renyi, Poisson arrivals.

The reference reads public datasets from disk instead (README.md:14-18); its
per-part loop (src/WithPartition/without_sorting_tc.cu:244-465) is what the
batched launch over these graphs replaces.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

RMAT_ABC = (0.57, 0.19, 0.19)


@dataclass(frozen=True)
class GraphSpec:
    """One synthetic graph: kind 'grid' (a=rows, b=cols, p), 'chung_lu' (a=n, gamma, mean, w_max),
    'er' (a=n, mean degree p, or m_raw when given), 'rmat' (a=scale, m_raw)."""
    name: str
    kind: str
    a: int
    b: int = 0
    p: float = 0.0
    gamma: float = 0.0
    mean: float = 0.0
    w_max: float = 0.0
    m_raw: int = 0
    seed: int = 1


def mixed_specs(seed: int = 1) -> list[GraphSpec]:
    """The 66 graphs of the mixed-shape batch, in launch order."""
    out = [GraphSpec("grid2000", "grid", 2000, 2000, p=0.3, seed=seed),
           GraphSpec("chung_lu_1M", "chung_lu", 1_000_000, gamma=2.1, mean=50.0, w_max=5e4, seed=seed)]
    rng = np.random.default_rng(seed)
    for i in range(64):
        nn = int(round(10 ** rng.uniform(3, 5)))
        out.append(GraphSpec(f"er{i}_{nn}", "er", nn, p=20.0, seed=seed * 1000 + i))
    return out


@dataclass(frozen=True)
class ServeRequest:
    kind: str      # "rmat" or "er"
    a: int         # RMAT scale, or ER vertex count
    m_raw: int     # raw edge samples
    seed: int
    arrive_s: float


def serving_stream(n_req: int = 10000, rate: float = 1000.0, seed: int = 1) -> list[ServeRequest]:
    """Request i: raw edges E_i log-uniform in [1e3, 5e6]; RMAT with probability 0.7
    (scale max(4, round(log2(E_i / 16))), edge factor E_i / 2^scale), else ER with
    n = max(16, E_i // 10) (mean degree about 20); seed seed * 1_000_003 + i; arrival
    times a Poisson process of ``rate`` requests per second."""
    rng = np.random.default_rng(seed)
    E = np.exp(rng.uniform(np.log(1e3), np.log(5e6), n_req)).astype(np.int64)
    is_rmat = rng.uniform(size=n_req) < 0.7
    arrive = np.cumsum(rng.exponential(1.0 / rate, n_req))
    out = []
    for i in range(n_req):
        e, s = int(E[i]), seed * 1_000_003 + i
        if is_rmat[i]:
            out.append(ServeRequest("rmat", max(4, int(round(np.log2(e / 16)))), e, s, float(arrive[i])))
        else:
            out.append(ServeRequest("er", max(16, e // 10), e, s, float(arrive[i])))
    return out


def generate(spec: GraphSpec, ws=None):
    """Raw edge samples of one spec on the current device: (n, src, dst)."""
    from . import core
    if spec.kind == "grid":
        return core.generate_grid(spec.a, spec.b, spec.p, seed=spec.seed)
    if spec.kind == "chung_lu":
        return core.generate_chung_lu(spec.a, spec.gamma, spec.mean, spec.w_max, seed=spec.seed, ws=ws)
    if spec.kind == "er":
        return core.generate_er(spec.a, spec.p, seed=spec.seed, m_raw=spec.m_raw or None)
    if spec.kind == "rmat":
        return core.generate_rmat(spec.a, 16, *RMAT_ABC, seed=spec.seed)
    raise ValueError(f"unknown graph kind {spec.kind}")


def mixed_graphs(seed: int = 1, ws=None, in_lists: bool = True) -> tuple[list, list[str]]:
    """The mixed-shape batch generated, cleaned and oriented on the device (with in-lists and
    segment words for split counting unless ``in_lists`` is False); returns (graphs, names)."""
    from . import core
    graphs, names = [], []
    for spec in mixed_specs(seed):
        n, s, d = generate(spec, ws)
        graphs.append(core.clean_orient(n, s, d, ws=ws, in_lists=in_lists))
        names.append(spec.name)
        del s, d
    return graphs, names
