"""Host-side API over the C ABI (device memory through torch, compute in liblstc).

The reference's host interface is a set of executables plus text files
(README.md:39-91): ``without_part_preprocess.py <file>`` (clean + orient),
``round_clean_sg <k> <part.csr> <node_parts>`` (halo extraction) and
``TC_128 <file.csr>`` / ``wo_sort_tc_256 <file> <k>`` (count).  The functions
here take the same roles on device-resident graphs:

    generate_*       synthetic raw edge samples in HBM (inputs the reference reads from disk)
    clean_orient     Med_dataset/without_part_preprocess.py:193-311
    partition_random Halo_creation_preprocess.py:288,325,362 (seeded)
    halo_extract     round_clean_subgraph.cu:357-902
    count            ShareKernelTCV62_update_v1.cu:43-296 / without_sorting_tc.cu:57-470

Errors raise LstcError (the reference prints and exits).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import torch

from ._lib import LstcError, Tune, check, load

__all__ = ["Graph", "generate_rmat", "generate_er", "generate_chung_lu", "generate_grid",
           "clean_orient", "count", "partition_random", "Workspace", "stream_handle"]


def _dev():
    if not torch.cuda.is_available():
        raise RuntimeError("trianglecount_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle(stream: torch.cuda.Stream | None = None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def _ptr(t: torch.Tensor | None):
    return None if t is None else C.c_void_p(t.data_ptr())


def empty_u32(n: int) -> torch.Tensor:
    return torch.empty(max(int(n), 1), dtype=torch.int32, device=_dev())


def empty_u64(n: int) -> torch.Tensor:
    return torch.empty(max(int(n), 1), dtype=torch.int64, device=_dev())


class Workspace:
    """Grow-only device scratch buffer handed to the *_workspace-sized calls."""

    def __init__(self):
        self.buf: torch.Tensor | None = None

    def get(self, nbytes: int) -> tuple[C.c_void_p, int]:
        nbytes = max(int(nbytes), 256)
        if self.buf is None or self.buf.numel() < nbytes:
            self.buf = torch.empty(nbytes, dtype=torch.uint8, device=_dev())
        return C.c_void_p(self.buf.data_ptr()), self.buf.numel()


_default_ws = Workspace()


@dataclass
class Graph:
    """Oriented CSR in HBM: uint64 row offsets ``off[n+1]`` and uint32 ids ``col[m]``.

    ``sym_off``/``sym_col`` optionally hold the symmetric (undirected) CSR that
    the halo extraction reads.  Tensors use int64/int32 storage for uint64/uint32.
    """
    n: int
    m: int
    off: torch.Tensor
    col: torch.Tensor
    sym_off: torch.Tensor | None = None
    sym_col: torch.Tensor | None = None
    n_owned: int | None = None
    meta: dict = field(default_factory=dict)

    @property
    def owned(self) -> int:
        return self.n if self.n_owned is None else self.n_owned

    @staticmethod
    def from_numpy(off, col, sym_off=None, sym_col=None, n_owned=None) -> "Graph":
        """Upload a host CSR (uint64 offsets, uint32 ids) into HBM."""
        import numpy as np
        dev = _dev()
        off = np.ascontiguousarray(off, dtype=np.uint64)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        n = len(off) - 1
        m = int(off[-1]) if n >= 0 and len(off) else 0

        def up(a, dt):
            a = a if len(a) else np.zeros(1, a.dtype)
            return torch.from_numpy(a.view(dt)).to(dev)

        g = Graph(n, m, up(off, np.int64), up(col, np.int32), n_owned=n_owned)
        if sym_off is not None:
            g.sym_off = up(np.ascontiguousarray(sym_off, np.uint64), np.int64)
            g.sym_col = up(np.ascontiguousarray(sym_col, np.uint32), np.int32)
        return g

    def to_numpy(self):
        import numpy as np
        off = self.off[: self.n + 1].cpu().numpy().view(np.uint64).copy()
        col = self.col[: self.m].cpu().numpy().view(np.uint32).copy()
        return off, col

    def sym_to_numpy(self):
        import numpy as np
        off = self.sym_off[: self.n + 1].cpu().numpy().view(np.uint64).copy()
        col = self.sym_col[: 2 * self.m].cpu().numpy().view(np.uint32).copy()
        return off, col


# ------------------------------------------------------------------ generators
def generate_rmat(scale: int, edge_factor: int = 16, a=0.57, b=0.19, c=0.19, seed: int = 1,
                  stream=None):
    """Raw RMAT samples (m_raw = edge_factor * 2**scale) -> (n, src, dst) in HBM."""
    m = int(edge_factor) << int(scale)
    src, dst = empty_u32(m), empty_u32(m)
    check(load().lstc_gen_rmat(scale, m, a, b, c, seed, _ptr(src), _ptr(dst),
                               C.c_void_p(stream_handle(stream))), "lstc_gen_rmat")
    return 1 << scale, src[:m], dst[:m]


def generate_er(n: int, mean_deg: float = 20.0, seed: int = 1, m_raw: int | None = None, stream=None):
    m = int(round(n * mean_deg / 2)) if m_raw is None else int(m_raw)
    src, dst = empty_u32(m), empty_u32(m)
    check(load().lstc_gen_er(n, m, seed, _ptr(src), _ptr(dst), C.c_void_p(stream_handle(stream))),
          "lstc_gen_er")
    return n, src[:m], dst[:m]


def generate_chung_lu(n: int, gamma=2.1, mean_deg=50.0, w_max=5e4, seed: int = 1, stream=None,
                      ws: Workspace | None = None):
    m = int(round(n * mean_deg / 2))
    src, dst = empty_u32(m), empty_u32(m)
    L = load()
    nb = C.c_size_t()
    check(L.lstc_gen_chung_lu_workspace(n, C.byref(nb)), "lstc_gen_chung_lu_workspace")
    wp, wb = (ws or _default_ws).get(nb.value)
    check(L.lstc_gen_chung_lu(n, gamma, mean_deg, w_max, m, seed, _ptr(src), _ptr(dst), wp, wb,
                              C.c_void_p(stream_handle(stream))), "lstc_gen_chung_lu")
    return n, src[:m], dst[:m]


def generate_grid(rows: int, cols: int, p_diag=0.3, seed: int = 1, stream=None):
    m = 3 * rows * cols
    src, dst = empty_u32(m), empty_u32(m)
    check(load().lstc_gen_grid(rows, cols, p_diag, seed, _ptr(src), _ptr(dst),
                               C.c_void_p(stream_handle(stream))), "lstc_gen_grid")
    return rows * cols, src[:m], dst[:m]


def partition_random(n: int, k: int, seed: int = 1, stream=None) -> torch.Tensor:
    part = empty_u32(n)
    check(load().lstc_partition_random(n, k, seed, _ptr(part), C.c_void_p(stream_handle(stream))),
          "lstc_partition_random")
    return part[:n]


# ------------------------------------------------------------------ (a) clean + orient
def clean_orient(n: int, src: torch.Tensor, dst: torch.Tensor, compact: bool = True,
                 keep_symmetric: bool = False, ws: Workspace | None = None, stream=None) -> Graph:
    """Drop loops/repeats, symmetrise, optionally drop isolated vertices, orient by (degree, id)."""
    L = load()
    m_raw = int(src.numel())
    if int(dst.numel()) != m_raw:
        raise ValueError("src and dst differ in length")
    nb = C.c_size_t()
    check(L.lstc_clean_orient_workspace(n, m_raw, C.byref(nb)), "lstc_clean_orient_workspace")
    wp, wb = (ws or _default_ws).get(nb.value)
    off = empty_u64(n + 1)
    col = empty_u32(m_raw)
    sym_off = empty_u64(n + 1) if keep_symmetric else None
    sym_col = empty_u32(2 * m_raw) if keep_symmetric else None
    n_out, m_out = C.c_uint32(), C.c_uint64()
    check(L.lstc_clean_orient(_ptr(src), _ptr(dst), m_raw, n, int(bool(compact)), _ptr(sym_off),
                              _ptr(sym_col), _ptr(off), _ptr(col), C.byref(n_out), C.byref(m_out),
                              wp, wb, C.c_void_p(stream_handle(stream))), "lstc_clean_orient")
    g = Graph(n_out.value, m_out.value, off, col, sym_off, sym_col)
    return g


# ------------------------------------------------------------------ (c) count
def count(g: Graph, n_owned: int | None = None, per_row: bool = False, tune: Tune | None = None,
          ws: Workspace | None = None, stream=None, out: torch.Tensor | None = None,
          sync: bool = True):
    """Exact count over owned rows; returns int (and the per-row tensor if asked).

    With ``sync=False`` the device result tensor is returned without waiting.
    """
    L = load()
    no = g.owned if n_owned is None else int(n_owned)
    nb = C.c_size_t()
    tp = C.byref(tune) if tune is not None else None
    check(L.lstc_count_workspace(g.n, no, g.m, tp, C.byref(nb)), "lstc_count_workspace")
    wp, wb = (ws or _default_ws).get(nb.value)
    res = out if out is not None else empty_u64(1)
    pr = empty_u64(no) if per_row else None
    check(L.lstc_count(_ptr(g.off), _ptr(g.col), g.n, no, g.m, _ptr(res), _ptr(pr), tp, wp, wb,
                       C.c_void_p(stream_handle(stream))), "lstc_count")
    if not sync:
        return (res, pr) if per_row else res
    total = int(res[0].item()) & 0xFFFFFFFFFFFFFFFF
    if per_row:
        import numpy as np
        return total, pr[:no].cpu().numpy().view(np.uint64)
    return total
