"""CPU oracle for the LSTC triangle-counting path (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only to check results.  The product
(``trianglecount_amd``) never imports it.

Thin numpy/ctypes wrapper over ``oracle/_build/liblstc_oracle.so`` (built from
``oracle/lstc_oracle.c`` by ``oracle/Makefile``).  Each wrapper names the C
function it calls; the C source cites the reference file:line it restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liblstc_oracle.so")
REF_UNDIR2DIR = os.path.join(HERE, "_ref", "undir2dir")

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


def build(ref: bool = False) -> None:
    """Compile the oracle (and, if asked and the reference tree exists, oracle/_ref)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if ref and os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.lo_mix64.restype = C.c_uint64
        L.lo_mix64.argtypes = [C.c_uint64]
        L.lo_stream_key.restype = C.c_uint64
        L.lo_stream_key.argtypes = [C.c_uint64, C.c_uint64]
        L.lo_rng64.restype = C.c_uint64
        L.lo_rng64.argtypes = [C.c_uint64, C.c_uint64]
        L.lo_gen_rmat.restype = None
        L.lo_gen_rmat.argtypes = [C.c_uint32, C.c_uint64, C.c_double, C.c_double, C.c_double,
                                  C.c_uint64, _u32p, _u32p]
        L.lo_gen_er.restype = None
        L.lo_gen_er.argtypes = [C.c_uint32, C.c_uint64, C.c_uint64, _u32p, _u32p]
        L.lo_gen_chung_lu.restype = C.c_int
        L.lo_gen_chung_lu.argtypes = [C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                      C.c_uint64, _u32p, _u32p]
        L.lo_gen_grid.restype = None
        L.lo_gen_grid.argtypes = [C.c_uint32, C.c_uint32, C.c_double, C.c_uint64, _u32p, _u32p]
        L.lo_grid_num_diagonals.restype = C.c_uint64
        L.lo_grid_num_diagonals.argtypes = [C.c_uint32, C.c_uint32, C.c_double, C.c_uint64]
        L.lo_clean.restype = None
        L.lo_clean.argtypes = [C.c_uint32, C.c_uint64, _u32p, _u32p, C.c_int,
                               C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), _u64p, _u32p]
        L.lo_orient.restype = None
        L.lo_orient.argtypes = [C.c_uint32, _u64p, _u32p, _u64p, _u32p]
        L.lo_count_merge.restype = C.c_uint64
        L.lo_count_merge.argtypes = [C.c_uint32, C.c_uint32, _u64p, _u32p, C.c_void_p, C.c_int]
        L.lo_count_merge_list.restype = C.c_uint64
        L.lo_count_merge_list.argtypes = [_u64p, _u32p, _u32p, C.c_uint64, C.c_int]
        L.lo_count_rows_k1.restype = C.c_uint64
        L.lo_count_rows_k1.argtypes = [C.c_uint32, C.c_uint32, _u64p, _u32p, C.c_void_p]
        L.lo_triangle_support.restype = None
        L.lo_triangle_support.argtypes = [C.c_uint32, C.c_uint32, _u64p, _u32p, C.c_void_p, C.c_void_p]
        L.lo_partition_random.restype = None
        L.lo_partition_random.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64, _u32p]
        L.lo_halo_extract.restype = None
        L.lo_halo_extract.argtypes = [C.c_uint32, _u64p, _u32p, _u32p, C.c_uint32, C.c_int,
                                      C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                      C.POINTER(C.c_uint64), C.c_void_p, C.c_void_p, C.c_void_p]
        L.lo_stats.restype = None
        L.lo_stats.argtypes = [C.c_uint32, _u64p, _u32p, C.c_void_p]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------- generators
def gen_rmat(scale: int, edge_factor: int = 16, a=0.57, b=0.19, c=0.19, seed: int = 1,
             m_raw: int | None = None):
    """lo_gen_rmat: raw RMAT samples (m_raw = edge_factor * 2**scale unless given)."""
    m = (int(edge_factor) << int(scale)) if m_raw is None else int(m_raw)
    s = np.empty(m, np.uint32)
    d = np.empty(m, np.uint32)
    lib().lo_gen_rmat(scale, m, a, b, c, seed, s, d)
    return 1 << scale, s, d


def gen_er(n: int, mean_deg: float = 20.0, seed: int = 1, m_raw: int | None = None):
    """lo_gen_er: m_raw = round(n * mean_deg / 2) uniform pairs."""
    m = int(round(n * mean_deg / 2)) if m_raw is None else int(m_raw)
    s = np.empty(m, np.uint32)
    d = np.empty(m, np.uint32)
    lib().lo_gen_er(n, m, seed, s, d)
    return n, s, d


def gen_chung_lu(n: int, gamma=2.1, mean_deg=50.0, w_max=5e4, seed: int = 1):
    m = int(round(n * mean_deg / 2))
    s = np.empty(m, np.uint32)
    d = np.empty(m, np.uint32)
    rc = lib().lo_gen_chung_lu(n, gamma, mean_deg, w_max, m, seed, s, d)
    if rc != 0:
        raise ValueError(f"lo_gen_chung_lu failed ({rc})")
    return n, s, d


def gen_grid(rows: int, cols: int, p_diag=0.3, seed: int = 1):
    m = 3 * rows * cols
    s = np.empty(m, np.uint32)
    d = np.empty(m, np.uint32)
    lib().lo_gen_grid(rows, cols, p_diag, seed, s, d)
    return rows * cols, s, d


def grid_num_diagonals(rows, cols, p_diag=0.3, seed=1) -> int:
    return int(lib().lo_grid_num_diagonals(rows, cols, p_diag, seed))


# ---------------------------------------------------------------- clean/orient
@dataclass
class CSR:
    n: int
    off: np.ndarray  # uint64 [n+1]
    col: np.ndarray  # uint32 [m]

    @property
    def m(self) -> int:
        return int(self.off[-1]) if len(self.off) else 0


def clean(n, src, dst, compact=True) -> CSR:
    """lo_clean: symmetric simple CSR (rows ascending)."""
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    m_raw = len(src)
    off = np.zeros(n + 1, np.uint64)
    col = np.empty(max(2 * m_raw, 1), np.uint32)
    n_out = C.c_uint32()
    m_out = C.c_uint64()
    lib().lo_clean(n, m_raw, src, dst, int(bool(compact)), C.byref(n_out), C.byref(m_out), off, col)
    nn = n_out.value
    return CSR(nn, off[: nn + 1].copy(), col[: 2 * m_out.value].copy())


def orient(sym: CSR) -> CSR:
    """lo_orient: keep s->d iff deg s < deg d or (equal and d < s)."""
    off = np.zeros(sym.n + 1, np.uint64)
    col = np.empty(max(sym.m // 2, 1), np.uint32)
    lib().lo_orient(sym.n, sym.off, sym.col, off, col)
    return CSR(sym.n, off, col[: int(off[-1])].copy())


def clean_orient(n, src, dst, compact=True):
    sym = clean(n, src, dst, compact)
    return sym, orient(sym)


# ---------------------------------------------------------------- counting
def count(g: CSR, n_owned: int | None = None, per_row: bool = False, threads: int = 0):
    """lo_count_merge over rows [0, n_owned)."""
    no = g.n if n_owned is None else int(n_owned)
    pr = np.zeros(max(no, 1), np.uint64) if per_row else None
    t = int(lib().lo_count_merge(g.n, no, g.off, g.col, _ptr(pr), int(threads)))
    return (t, pr[:no]) if per_row else t


def count_rows(off, col, rows, threads: int = 0) -> int:
    """lo_count_merge_list: merge counter over the listed rows only."""
    rows = np.ascontiguousarray(rows, np.uint32)
    return int(lib().lo_count_merge_list(np.ascontiguousarray(off, np.uint64),
                                         np.ascontiguousarray(col, np.uint32), rows, len(rows),
                                         int(threads)))


def count_k1(g: CSR, n_owned: int | None = None):
    """lo_count_rows_k1: literal block-per-row restatement; returns (total, g_sum)."""
    no = g.n if n_owned is None else int(n_owned)
    gs = np.zeros(max(no, 1), np.uint64)
    t = int(lib().lo_count_rows_k1(g.n, no, g.off, g.col, _ptr(gs)))
    return t, gs[:no]


def triangle_support(g: CSR, n_owned: int | None = None):
    """lo_triangle_support: (support[m] uint32 by arc position, per_vertex[n] uint64)."""
    no = g.n if n_owned is None else int(n_owned)
    sup = np.zeros(max(g.m, 1), np.uint32)
    pv = np.zeros(max(g.n, 1), np.uint64)
    lib().lo_triangle_support(g.n, no, g.off, g.col, _ptr(sup), _ptr(pv))
    return sup[: g.m], pv[: g.n]


# ---------------------------------------------------------------- partition/halo
def partition_random(n, k, seed=1):
    part = np.empty(max(n, 1), np.uint32)
    lib().lo_partition_random(n, k, seed, part)
    return part[:n]


@dataclass
class Halo:
    csr: CSR
    n_owned: int
    l2g: np.ndarray


def halo_extract(sym: CSR, part, p, mode=0) -> Halo:
    part = np.ascontiguousarray(part, np.uint32)
    nl, no, ml = C.c_uint32(), C.c_uint32(), C.c_uint64()
    L = lib()
    L.lo_halo_extract(sym.n, sym.off, sym.col, part, p, mode, C.byref(nl), C.byref(no),
                      C.byref(ml), None, None, None)
    off = np.zeros(nl.value + 1, np.uint64)
    col = np.empty(max(ml.value, 1), np.uint32)
    l2g = np.empty(max(nl.value, 1), np.uint32)
    L.lo_halo_extract(sym.n, sym.off, sym.col, part, p, mode, C.byref(nl), C.byref(no),
                      C.byref(ml), _ptr(off), _ptr(col), _ptr(l2g))
    return Halo(CSR(nl.value, off, col[: ml.value].copy()), no.value, l2g[: nl.value].copy())


# ---------------------------------------------------------------- stats
class _Stats(C.Structure):
    _fields_ = [("n", C.c_uint64), ("m", C.c_uint64), ("max_out", C.c_uint64),
                ("sum_out_sq", C.c_uint64), ("sum_in_out", C.c_uint64), ("rows_ge2", C.c_uint64)]


def stats(g: CSR) -> dict:
    st = _Stats()
    lib().lo_stats(g.n, g.off, g.col, C.byref(st))
    d = {k: int(getattr(st, k)) for k, _ in _Stats._fields_}
    return d


def b_alg(st: dict) -> int:
    """Merge-model algorithmic bytes: 8(n+1) + 4m + sum_{(u,v)} [16 + 4(d+(u)+d+(v))]."""
    return 8 * (st["n"] + 1) + 4 * st["m"] + 16 * st["m"] + 4 * (st["sum_out_sq"] + st["sum_in_out"])


# ---------------------------------------------------------------- reference tool
def ref_undir2dir_available() -> bool:
    return os.path.exists(REF_UNDIR2DIR)


def run_ref_undir2dir(sym: CSR, workdir: str) -> CSR:
    """Orient ``sym`` with the reference's compiled preprocess_CPU/Undir2Dir.c.

    Input format (Undir2Dir.c:31-94): in_deg count, in_deg[], then one part with
    org_id_s row_ptr_s col_idx_s t_ver, org_id[], row_ptr[], col_idx[].
    Output (Undir2Dir.c:137-150): t_ver row_ptr_s pos, then row_ptr_Dir[], col[].
    """
    deg = np.diff(sym.off).astype(np.uint64)
    inp = os.path.join(workdir, "undir.txt")
    out = os.path.join(workdir, "dir.txt")
    with open(inp, "w") as f:
        f.write(f"{sym.n}\n" + " ".join(str(int(x)) for x in deg) + "\n")
        f.write(f"{sym.n} {sym.n + 1} {sym.m} {sym.n}\n")
        f.write(" ".join(str(i) for i in range(sym.n)) + "\n")
        f.write(" ".join(str(int(x)) for x in sym.off) + "\n")
        f.write(" ".join(str(int(x)) for x in sym.col) + "\n")
    subprocess.run([REF_UNDIR2DIR, inp, out, "1"], check=True, stdout=subprocess.DEVNULL,
                   stderr=subprocess.DEVNULL, cwd=workdir)
    toks = open(out).read().split()
    t_ver, rps, pos = int(toks[0]), int(toks[1]), int(toks[2])
    vals = np.array([int(x) for x in toks[3:]], dtype=np.uint64)
    off = vals[:rps]
    col = vals[rps: rps + pos].astype(np.uint32)
    assert t_ver == sym.n
    return CSR(sym.n, off.astype(np.uint64), col)
