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
           "clean_orient", "count", "partition_random", "halo_extract", "Workspace", "stream_handle"]


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
    in_off: torch.Tensor | None = None  # target-side in-lists (split count), optional
    in_col: torch.Tensor | None = None
    in_owned: int | None = None  # the n_owned the in-lists were built for (owned sources only)
    col_seg: torch.Tensor | None = None  # segment words of col / in_col (lstc_segments), optional
    in_seg: torch.Tensor | None = None
    scol: torch.Tensor | None = None  # streaming copy the segment words point into (rows 128-byte aligned)

    def split_ready(self, n_owned: int | None = None) -> bool:
        """True when the in-lists fit a count over the first ``n_owned`` rows."""
        no = self.owned if n_owned is None else int(n_owned)
        return self.in_off is not None and self.in_owned == no

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


def partition_contig(n: int, k: int, stream=None) -> torch.Tensor:
    """Contiguous blocks of n // k vertices, the remainder in part 0 (lstc_partition_contig)."""
    part = empty_u32(n)
    check(load().lstc_partition_contig(n, k, _ptr(part), C.c_void_p(stream_handle(stream))), "lstc_partition_contig")
    return part[:n]


def partition_from_file(path: str, n: int, k: int, stream=None) -> torch.Tensor:
    """An externally computed assignment (e.g. METIS) from a node-parts file (lstc_partition_read)."""
    part = empty_u32(n)
    check(load().lstc_partition_read(str(path).encode(), n, k, _ptr(part), C.c_void_p(stream_handle(stream))),
          "lstc_partition_read")
    return part[:n]


# ------------------------------------------------------------------ (a) clean + orient
def clean_orient(n: int, src: torch.Tensor, dst: torch.Tensor, compact: bool = True,
                 keep_symmetric: bool = False, in_lists: bool = False, ws: Workspace | None = None,
                 stream=None, stream_copy: bool = False) -> Graph:
    """Drop loops/repeats, symmetrise, optionally drop isolated vertices, orient by (degree, id).

    ``in_lists`` also emits the target-side in-lists used by the split count; with
    ``stream_copy`` also a copy of the columns whose rows start every 128 bytes, which
    the count then streams its lists from (lstc_clean_orient_stream).
    """
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
    in_off = empty_u64(n + 1) if in_lists else None
    in_col = empty_u32(m_raw) if in_lists else None
    n_out, m_out = C.c_uint32(), C.c_uint64()
    if stream_copy and not in_lists:
        raise ValueError("clean_orient: a streaming copy comes with the in-lists (in_lists=True)")
    if in_lists:  # in-lists and segment words in the same pass (lstc_clean_orient_split)
        col_seg, in_seg, n_in = empty_u64(m_raw), empty_u64(m_raw), C.c_uint64()
        sh = C.c_void_p(stream_handle(stream))
        scol = None
        if stream_copy:
            scol = empty_u32(max(int(L.lstc_stream_copy_capacity(n, m_raw)), 1))
            s_used = C.c_uint64()
            check(L.lstc_clean_orient_stream(_ptr(src), _ptr(dst), m_raw, n, int(bool(compact)), _ptr(sym_off),
                                             _ptr(sym_col), _ptr(off), _ptr(col), _ptr(in_off), _ptr(in_col),
                                             _ptr(col_seg), _ptr(in_seg), _ptr(scol), C.byref(n_out), C.byref(m_out),
                                             C.byref(n_in), C.byref(s_used), wp, wb, sh), "lstc_clean_orient_stream")
            scol = scol[: max(s_used.value, 1)]
        else:
            check(L.lstc_clean_orient_split(_ptr(src), _ptr(dst), m_raw, n, int(bool(compact)), _ptr(sym_off),
                                            _ptr(sym_col), _ptr(off), _ptr(col), _ptr(in_off), _ptr(in_col),
                                            _ptr(col_seg), _ptr(in_seg), C.byref(n_out), C.byref(m_out),
                                            C.byref(n_in), wp, wb, sh), "lstc_clean_orient_split")
        g = Graph(n_out.value, m_out.value, off, col, sym_off, sym_col)
        g.in_off, g.in_col, g.in_owned = in_off, in_col, g.n
        g.col_seg, g.in_seg, g.scol = col_seg, in_seg, scol
        g.meta["in_entries"] = n_in.value
        return g
    check(L.lstc_clean_orient(_ptr(src), _ptr(dst), m_raw, n, int(bool(compact)), _ptr(sym_off),
                              _ptr(sym_col), _ptr(off), _ptr(col), _ptr(in_off), _ptr(in_col),
                              C.byref(n_out), C.byref(m_out), wp, wb, C.c_void_p(stream_handle(stream))),
          "lstc_clean_orient")
    return Graph(n_out.value, m_out.value, off, col, sym_off, sym_col)


def segments(g: Graph, stream=None) -> Graph:
    """Attach the segment words of ``g``'s col and in-lists (lstc_segments); needs the in-lists."""
    if g.in_off is None:
        raise ValueError("segments: build the in-lists first (clean_orient(in_lists=True) or in_lists(g))")
    L = load()
    n_in = int(g.in_off[g.n].item())
    col_seg = empty_u64(max(g.m, 1))
    in_seg = empty_u64(max(n_in, 1))
    check(L.lstc_segments(_ptr(g.off), _ptr(g.col), g.n, g.m, _ptr(g.in_col), n_in, _ptr(col_seg), _ptr(in_seg),
                          C.c_void_p(stream_handle(stream))), "lstc_segments")
    g.col_seg, g.in_seg, g.scol = col_seg, in_seg, None  # these words point into col
    return g


def stream_copy(g: Graph, ws: Workspace | None = None, stream=None) -> Graph:
    """Attach a row-aligned streaming copy of ``g``'s columns and point its segment words into
    it (lstc_stream_copy), so its split counts read every list from the start of a cache line.
    ``g`` needs its in-lists and segment words (in_lists(g) or clean_orient(in_lists=True))."""
    if g.col_seg is None or g.in_off is None:
        raise ValueError("stream_copy: build the in-lists and segment words first (in_lists(g))")
    if g.scol is not None:
        return g
    L = load()
    nb = C.c_size_t()
    check(L.lstc_stream_copy_workspace(g.n, C.byref(nb)), "lstc_stream_copy_workspace")
    wp, wb = (ws or _default_ws).get(nb.value)
    no = g.in_owned if g.in_owned is not None else g.owned
    n_words = g.m if no == g.n else int(g.off[no].item())
    n_in = int(g.meta.get("in_entries", int(g.in_off[g.n].item())))
    cap = g.m + 31 * g.n
    scol = empty_u32(max(cap, 1))
    used = C.c_uint64()
    check(L.lstc_stream_copy(_ptr(g.off), _ptr(g.col), g.n, g.m, n_words, _ptr(g.col_seg), _ptr(g.in_col), n_in,
                             _ptr(g.in_seg), _ptr(scol), cap, C.byref(used), wp, wb,
                             C.c_void_p(stream_handle(stream))), "lstc_stream_copy")
    g.scol = scol[: max(used.value, 1)]
    return g


def in_lists(g: Graph, ws: Workspace | None = None, stream=None, segment_words: bool = True) -> Graph:
    """Attach target-side in-lists (owned sources only) so ``g`` is counted split: with their
    segment words and those of col in the same passes (lstc_in_lists_segments), or the in-lists
    alone (lstc_in_lists)."""
    L = load()
    no = g.owned
    nb = C.c_size_t()
    check(L.lstc_in_lists_workspace(g.n, no, g.m, C.byref(nb)), "lstc_in_lists_workspace")
    wp, wb = (ws or _default_ws).get(nb.value)
    # in-list entries are arcs of owned rows, which come first: off[n_owned] of them at most (m for a
    # whole graph), and the column words the count reads are theirs too
    mo = g.m if no == g.n else int(g.off[no].item())
    in_off = empty_u64(g.n + 1)
    in_col = empty_u32(max(mo, 1))
    k = C.c_uint64()
    sh = C.c_void_p(stream_handle(stream))
    if segment_words:
        col_seg = empty_u64(max(mo, 1))
        in_seg = empty_u64(max(mo, 1))
        check(L.lstc_in_lists_segments(_ptr(g.off), _ptr(g.col), g.n, no, g.m, _ptr(in_off), _ptr(in_col),
                                       _ptr(col_seg), _ptr(in_seg), C.byref(k), wp, wb, sh), "lstc_in_lists_segments")
        g.col_seg, g.in_seg, g.scol = col_seg, in_seg[: max(k.value, 1)], None
    else:
        check(L.lstc_in_lists(_ptr(g.off), _ptr(g.col), g.n, no, g.m, _ptr(in_off), _ptr(in_col), C.byref(k), wp,
                              wb, sh), "lstc_in_lists")
        g.col_seg = g.in_seg = g.scol = None
    g.in_off, g.in_col, g.in_owned = in_off, in_col, no
    g.meta["in_entries"] = k.value
    return g


# ------------------------------------------------------------------ (c) count
def count(g: Graph, n_owned: int | None = None, per_row: bool = False, tune: Tune | None = None,
          ws: Workspace | None = None, stream=None, out: torch.Tensor | None = None,
          sync: bool = True, split: bool = True):
    """Exact count over owned rows; returns int (and the per-row tensor if asked).

    A graph carrying in-lists built for this ``n_owned`` is counted with
    lstc_count_split unless ``split`` is False.  With ``sync=False`` the device
    result tensor is returned without waiting.
    """
    L = load()
    no = g.owned if n_owned is None else int(n_owned)
    nb = C.c_size_t()
    tp = C.byref(tune) if tune is not None else None
    check(L.lstc_count_workspace(g.n, no, g.m, tp, C.byref(nb)), "lstc_count_workspace")
    wp, wb = (ws or _default_ws).get(nb.value)
    res = out if out is not None else empty_u64(1)
    pr = empty_u64(no) if per_row else None
    if g.split_ready(no) and not per_row and split and g.scol is not None and g.col_seg is not None:
        check(L.lstc_count_split_stream(_ptr(g.off), _ptr(g.col), _ptr(g.in_off), _ptr(g.in_col), _ptr(g.col_seg),
                                        _ptr(g.in_seg), _ptr(g.scol), int(g.scol.numel()), g.n, no, g.m, _ptr(res),
                                        tp, wp, wb, C.c_void_p(stream_handle(stream))), "lstc_count_split_stream")
    elif g.split_ready(no) and not per_row and split:
        check(L.lstc_count_split(_ptr(g.off), _ptr(g.col), _ptr(g.in_off), _ptr(g.in_col), _ptr(g.col_seg),
                                 _ptr(g.in_seg), g.n, no, g.m, _ptr(res), tp, wp, wb,
                                 C.c_void_p(stream_handle(stream))), "lstc_count_split")
    else:
        check(L.lstc_count(_ptr(g.off), _ptr(g.col), g.n, no, g.m, _ptr(res), _ptr(pr), tp, wp, wb,
                           C.c_void_p(stream_handle(stream))), "lstc_count")
    if not sync:
        return (res, pr) if per_row else res
    total = int(res[0].item()) & 0xFFFFFFFFFFFFFFFF
    if per_row:
        import numpy as np
        return total, pr[:no].cpu().numpy().view(np.uint64)
    return total


def triangle_support(g: Graph, n_owned: int | None = None, per_edge: bool = True, per_vertex: bool = True,
                     split: bool | None = None, ws: Workspace | None = None, stream=None):
    """Per-edge support (uint32 per arc of g.col) and per-vertex triangle counts (uint64 per row),
    over the triangles whose source row is owned.  With the graph's in-lists (split, the default
    when it has them) lstc_triangle_support_split, else lstc_triangle_support.  Returns device
    tensors (support, per_vertex), either None when not asked for."""
    no = g.owned if n_owned is None else int(n_owned)
    if split is None:
        split = g.split_ready(no)
    if split and not g.split_ready(no):
        raise ValueError("split support needs in-lists of the same owned rows: clean_orient(..., in_lists=True) "
                         "or in_lists(g)")
    sup = empty_u32(g.m) if per_edge else None
    pv = empty_u64(g.n) if per_vertex else None
    L = load()
    sh = C.c_void_p(stream_handle(stream))
    if split:
        nb = C.c_size_t()
        check(L.lstc_triangle_support_split_workspace(g.n, g.m, C.byref(nb)), "lstc_triangle_support_split_workspace")
        wp, wb = (ws or _default_ws).get(nb.value)
        check(L.lstc_triangle_support_split(_ptr(g.off), _ptr(g.col), _ptr(g.in_off), _ptr(g.in_col), g.n, no, g.m,
                                            _ptr(sup), _ptr(pv), wp, wb, sh), "lstc_triangle_support_split")
    else:
        check(L.lstc_triangle_support(_ptr(g.off), _ptr(g.col), g.n, no, g.m, _ptr(sup), _ptr(pv), sh),
              "lstc_triangle_support")
    return (sup[: g.m] if sup is not None else None), (pv[: g.n] if pv is not None else None)


# ------------------------------------------------------------------ (b) halo
def halo_extract(g: Graph, part: torch.Tensor, p: int, mode: int = 0, with_l2g: bool = False,
                 ws: Workspace | None = None, stream=None, split: bool = False) -> Graph:
    """Halo subgraph of part ``p`` from the symmetric CSR kept in ``g`` (keep_symmetric=True).

    Returns an oriented CSR whose first ``n_owned`` rows are the owned vertices;
    ``meta['l2g']`` holds the local -> global map when asked.  ``split`` also
    builds its in-lists (lstc_in_lists) so the part is counted split.
    """
    if g.sym_off is None:
        raise ValueError("halo_extract needs the symmetric CSR: clean_orient(..., keep_symmetric=True)")
    L = load()
    nb = C.c_size_t()
    check(L.lstc_halo_workspace(g.n, C.byref(nb)), "lstc_halo_workspace")
    wp, wb = (ws or _default_ws).get(nb.value)
    sh = C.c_void_p(stream_handle(stream))
    nl, no, ml = C.c_uint32(), C.c_uint32(), C.c_uint64()
    check(L.lstc_halo_sizes(_ptr(g.sym_off), _ptr(g.sym_col), g.n, _ptr(part), p, mode, C.byref(nl),
                            C.byref(no), C.byref(ml), wp, wb, sh), "lstc_halo_sizes")
    off = empty_u64(nl.value + 1)
    col = empty_u32(ml.value)
    l2g = empty_u32(nl.value) if with_l2g else None
    check(L.lstc_halo_extract(_ptr(g.sym_off), _ptr(g.sym_col), g.n, _ptr(part), p, mode, _ptr(off),
                              _ptr(col), _ptr(l2g), nl.value, ml.value, wp, wb, sh), "lstc_halo_extract")
    h = Graph(nl.value, ml.value, off, col, n_owned=no.value)
    if with_l2g:
        h.meta["l2g"] = l2g[: nl.value]
    if split:
        in_lists(h, ws=ws, stream=stream)
    return h


def halo_extract_parts(g: Graph, part: torch.Tensor, k: int, mode: int = 0, with_l2g: bool = False,
                       ws: Workspace | None = None, stream=None, split: bool = False,
                       one_pass_limit_bytes: int = 8 << 30) -> list:
    """Halo subgraphs of all ``k`` parts: the same graphs as
    ``[halo_extract(g, part, p, mode) for p in range(k)]``, built by passes that serve every part.

    When k upper-bound buffers (n_local <= n, m_local <= the undirected edges) take at most
    ``one_pass_limit_bytes``, one call sizes and fills them (lstc_halo_parts_build) and the parts
    are views of those buffers; otherwise the sizes come first (lstc_halo_parts_sizes) and the
    buffers are allocated exactly (lstc_halo_parts_extract)."""
    from ._lib import HaloOut
    if g.sym_off is None:
        raise ValueError("halo_extract_parts needs the symmetric CSR: clean_orient(..., keep_symmetric=True)")
    L = load()
    k = int(k)
    nb = C.c_size_t()
    check(L.lstc_halo_parts_workspace(g.n, k, C.byref(nb)), "lstc_halo_parts_workspace")
    wp, wb = (ws or _default_ws).get(nb.value)
    sh = C.c_void_p(stream_handle(stream))
    nl, no, ml = (C.c_uint32 * k)(), (C.c_uint32 * k)(), (C.c_uint64 * k)()
    one_pass = k * (8 * (g.n + 1) + 4 * g.m + (4 * g.n if with_l2g else 0)) <= one_pass_limit_bytes
    if one_pass:
        caps = [(g.n, g.m)] * k
    else:
        check(L.lstc_halo_parts_sizes(_ptr(g.sym_off), _ptr(g.sym_col), g.n, _ptr(part), k, mode, nl, no, ml, wp,
                                      wb, sh), "lstc_halo_parts_sizes")
        caps = [(nl[p], ml[p]) for p in range(k)]
    outs = (HaloOut * k)()
    bufs = []
    for p, (cn, cm) in enumerate(caps):
        off, col = empty_u64(cn + 1), empty_u32(cm)
        l2g = empty_u32(cn) if with_l2g else None
        outs[p].off, outs[p].col = off.data_ptr(), col.data_ptr()
        outs[p].l2g = l2g.data_ptr() if with_l2g else None
        outs[p].n_local_cap, outs[p].m_local_cap = cn, cm
        bufs.append((off, col, l2g))
    if one_pass:
        check(L.lstc_halo_parts_build(_ptr(g.sym_off), _ptr(g.sym_col), g.n, _ptr(part), k, mode, outs, nl, no, ml,
                                      wp, wb, sh), "lstc_halo_parts_build")
    else:
        check(L.lstc_halo_parts_extract(_ptr(g.sym_off), _ptr(g.sym_col), g.n, _ptr(part), k, mode, outs, wp, wb,
                                        sh), "lstc_halo_parts_extract")
    parts = []
    for p, (off, col, l2g) in enumerate(bufs):
        h = Graph(nl[p], ml[p], off[: nl[p] + 1], col[: max(ml[p], 1)], n_owned=no[p])
        if with_l2g:
            h.meta["l2g"] = l2g[: nl[p]]
        if split:
            in_lists(h, ws=ws, stream=stream)
        parts.append(h)
    return parts


# ------------------------------------------------------------------ batching
def _descs(graphs, per_row=None):
    from ._lib import GraphDesc
    arr = (GraphDesc * len(graphs))()
    for i, g in enumerate(graphs):
        arr[i].off = g.off.data_ptr()
        arr[i].col = g.col.data_ptr()
        arr[i].per_row = per_row[i].data_ptr() if per_row is not None else None
        if g.split_ready():
            arr[i].in_off = g.in_off.data_ptr()
            arr[i].in_col = g.in_col.data_ptr()
            if g.col_seg is not None:
                arr[i].col_seg = g.col_seg.data_ptr()
                arr[i].in_seg = g.in_seg.data_ptr()
                if g.scol is not None:  # the words point into the streaming copy
                    arr[i].scol = g.scol.data_ptr()
                    arr[i].scol_len = int(g.scol.numel())
        arr[i].m = g.m
        arr[i].n_rows = g.n
        arr[i].n_owned = g.owned
    return arr


def count_batch(graphs: list, tune: Tune | None = None, ws: Workspace | None = None, stream=None,
                out: torch.Tensor | None = None, sync: bool = True):
    """All graphs in one persistent launch; returns a list of per-graph counts."""
    L = load()
    G = len(graphs)
    arr = _descs(graphs)
    tp = C.byref(tune) if tune is not None else None
    nb = C.c_size_t()
    check(L.lstc_count_batch_workspace(arr, G, tp, C.byref(nb)), "lstc_count_batch_workspace")
    wp, wb = (ws or _default_ws).get(nb.value)
    res = out if out is not None else empty_u64(G)
    check(L.lstc_count_batch(arr, G, _ptr(res), tp, wp, wb, C.c_void_p(stream_handle(stream))),
          "lstc_count_batch")
    if not sync:
        return res
    return [int(x) & 0xFFFFFFFFFFFFFFFF for x in res[:G].cpu().tolist()]


class BatchPlan:
    """Resident batch plan: descriptors + workspace on the device, optional hipGraph replay."""

    def __init__(self, graphs: list, tune: Tune | None = None):
        self.L = load()
        self.graphs = list(graphs)  # keep the tensors alive
        self.G = len(graphs)
        self._arr = _descs(graphs)
        self.h = C.c_void_p()
        check(self.L.lstc_batch_plan_create(self._arr, self.G, C.byref(tune) if tune is not None else None,
                                            C.byref(self.h)), "lstc_batch_plan_create")
        self.counts = empty_u64(self.G)

    def run(self, stream=None):
        check(self.L.lstc_batch_plan_run(self.h, _ptr(self.counts), C.c_void_p(stream_handle(stream))),
              "lstc_batch_plan_run")
        return self.counts

    def launch_graph(self, stream=None):
        check(self.L.lstc_batch_plan_launch_graph(self.h, _ptr(self.counts), C.c_void_p(stream_handle(stream))),
              "lstc_batch_plan_launch_graph")
        return self.counts

    def results(self):
        torch.cuda.synchronize()
        return [int(x) & 0xFFFFFFFFFFFFFFFF for x in self.counts[: self.G].cpu().tolist()]

    def close(self):
        if self.h:
            self.L.lstc_batch_plan_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ serving
class Server:
    """Continuously batched count server (lstc_server_*) with an HBM budget."""

    def __init__(self, hbm_budget_bytes: int, max_batch: int = 1024, tune: Tune | None = None):
        from ._lib import Request, Result, ServerStats  # noqa: F401
        self.L = load()
        _dev()
        self.h = C.c_void_p()
        check(self.L.lstc_server_create(int(hbm_budget_bytes), int(max_batch),
                                        C.byref(tune) if tune is not None else None, C.byref(self.h)),
              "lstc_server_create")
        self._keep = {}

    def submit_rmat(self, scale: int, m_raw: int, seed: int, a=0.57, b=0.19, c=0.19) -> int:
        from ._lib import LSTC_REQ_RMAT, Request
        r = Request(kind=LSTC_REQ_RMAT, scale=scale, m_raw=m_raw, seed=seed, a=a, b=b, c=c)
        return self._submit(r)

    def submit_er(self, n: int, m_raw: int, seed: int) -> int:
        from ._lib import LSTC_REQ_ER, Request
        return self._submit(Request(kind=LSTC_REQ_ER, n=n, m_raw=m_raw, seed=seed))

    def submit_graph(self, g: "Graph") -> int:
        from ._lib import LSTC_REQ_CSR, Request
        r = Request(kind=LSTC_REQ_CSR, n_rows=g.n, off=g.off.data_ptr(), col=g.col.data_ptr(), m=g.m)
        rid = self._submit(r)
        self._keep[rid] = g  # the device CSR must outlive the request
        return rid

    def _submit(self, r) -> int:
        rid = C.c_uint64()
        check(self.L.lstc_server_submit(self.h, C.byref(r), C.byref(rid)), "lstc_server_submit")
        return rid.value

    def poll(self, rid: int):
        from ._lib import LSTC_ERR_NOT_READY, Result
        res = Result()
        st = self.L.lstc_server_poll(self.h, rid, C.byref(res))
        if st == LSTC_ERR_NOT_READY:
            return None
        check(st, "lstc_server_poll")
        self._keep.pop(rid, None)
        return {"count": int(res.count), "edges": int(res.edges), "latency_s": float(res.latency_s)}

    def wait(self, rid: int, timeout_s: float = 600.0):
        from ._lib import Result
        res = Result()
        check(self.L.lstc_server_wait(self.h, rid, float(timeout_s), C.byref(res)), "lstc_server_wait")
        self._keep.pop(rid, None)
        return {"count": int(res.count), "edges": int(res.edges), "latency_s": float(res.latency_s)}

    def cancel(self, rid: int) -> bool:
        """Cancels a request not admitted yet (its wait then raises LstcError with status
        LSTC_ERR_CANCELLED); False when it is already admitted or finished (it keeps its result)."""
        from ._lib import LSTC_ERR_NOT_READY
        st = self.L.lstc_server_cancel(self.h, rid)
        if st == LSTC_ERR_NOT_READY:
            return False
        check(st, "lstc_server_cancel")
        return True

    def stats(self) -> dict:
        from ._lib import ServerStats
        st = ServerStats()
        check(self.L.lstc_server_get_stats(self.h, C.byref(st)), "lstc_server_get_stats")
        return st.as_dict()

    def close(self):
        if self.h:
            self.L.lstc_server_destroy(self.h)
            self.h = C.c_void_p()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ diagnostics
def selftest() -> int:
    """Run lstc_selftest on the current device; returns the number of wrong lanes (0 = pass)."""
    _dev()
    bad = C.c_uint32()
    check(load().lstc_selftest(C.byref(bad)), "lstc_selftest")
    return bad.value


# ------------------------------------------------------------------ autotuning
def shape_class(graphs: list) -> int:
    sc = C.c_uint32()
    check(load().lstc_shape_class(_descs(graphs), len(graphs), C.byref(sc)), "lstc_shape_class")
    return sc.value


def work_stats(graphs: list, tune: Tune | None = None, ws: Workspace | None = None, stream=None) -> list[dict]:
    """Per-graph work of one count and its algorithmic bytes (lstc_work_stats)."""
    from ._lib import WorkStats
    L = load()
    G = len(graphs)
    nb = C.c_size_t()
    check(L.lstc_work_stats_workspace(G, C.byref(nb)), "lstc_work_stats_workspace")
    wp, wb = (ws or _default_ws).get(nb.value)
    out = (WorkStats * max(G, 1))()
    check(L.lstc_work_stats(_descs(graphs), G, C.byref(tune) if tune is not None else None, out, wp, wb,
                            C.c_void_p(stream_handle(stream))), "lstc_work_stats")
    return [out[i].as_dict() for i in range(G)]


class KernelTimer:
    """Count-kernel time from hipEvents the library records around each launch (lstc_timing_*).

        with KernelTimer() as t:
            ... counts ...
        t.ms, t.launches      # summed kernel milliseconds and launches inside the block
    """

    def __enter__(self):
        L = load()
        check(L.lstc_timing_read(None, None), "lstc_timing_read")  # drop earlier records
        check(L.lstc_timing_enable(1), "lstc_timing_enable")
        self.ms, self.launches = 0.0, 0
        return self

    def __exit__(self, *exc):
        L = load()
        check(L.lstc_timing_enable(0), "lstc_timing_enable")
        ms, n = C.c_double(), C.c_uint64()
        check(L.lstc_timing_read(C.byref(ms), C.byref(n)), "lstc_timing_read")
        self.ms, self.launches = ms.value, n.value
        return False


def tune_resolve(graphs: list, tune: Tune | None = None) -> Tune:
    """The tuning a launch over ``graphs`` runs with (lstc_tune_resolve; host only)."""
    out = Tune()
    check(load().lstc_tune_resolve(_descs(graphs), len(graphs), C.byref(tune) if tune is not None else None,
                                   C.byref(out)), "lstc_tune_resolve")
    return out


def autotune(graphs: list, reps: int = 3, cache: str | None = None, stream=None):
    """Pick lstc_tune for this batch (cached per device and shape class when ``cache`` is a path).

    Returns (tune, ms_per_launch or None when served from the cache, from_cache).
    """
    L = load()
    sc = shape_class(graphs)
    t = Tune()
    if cache is not None:
        if L.lstc_tune_cache_lookup(cache.encode(), sc, C.byref(t)) == 0:
            return t, None, True
    ms = C.c_double()
    check(L.lstc_autotune(_descs(graphs), len(graphs), reps, C.byref(t), C.byref(ms),
                          C.c_void_p(stream_handle(stream))), "lstc_autotune")
    if cache is not None:
        check(L.lstc_tune_cache_store(cache.encode(), sc, C.byref(t), ms.value), "lstc_tune_cache_store")
    return t, ms.value, False
