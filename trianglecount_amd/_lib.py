"""ctypes binding of liblstc.so (the C ABI declared in include/lstc.h).

The library is built in-tree by ``python -m trianglecount_amd.build``.  Loading
fails loudly when it is missing: there is no CPU fallback on the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import re

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "_lib", "liblstc.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "lstc.h")

LSTC_OK = 0
LSTC_ERR_INVALID = -1
LSTC_ERR_OOM = -2
LSTC_ERR_OVER_BUDGET = -3
LSTC_ERR_HIP = -4
LSTC_ERR_WORKSPACE = -5
LSTC_ERR_NOT_READY = -6
LSTC_ERR_CANCELLED = -7
LSTC_ERR_DEVICE = -8


class LstcError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"lstc status {status}: {msg}")
        self.status = status


class Tune(C.Structure):
    """lstc_tune (include/lstc.h)."""
    _fields_ = [("table_scale", C.c_uint32), ("small_max", C.c_uint32),
                ("small_group", C.c_uint32), ("grid_per_cu", C.c_uint32),
                ("unroll", C.c_uint32), ("tiny_max", C.c_uint32), ("waves", C.c_uint32),
                ("reserved", C.c_uint32 * 1)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_ if k != "reserved"}


class GraphDesc(C.Structure):
    """lstc_graph_desc (include/lstc.h)."""
    _fields_ = [("off", C.c_void_p), ("col", C.c_void_p), ("per_row", C.c_void_p),
                ("in_off", C.c_void_p), ("in_col", C.c_void_p),
                ("m", C.c_uint64), ("n_rows", C.c_uint32), ("n_owned", C.c_uint32),
                ("col_seg", C.c_void_p), ("in_seg", C.c_void_p), ("scol", C.c_void_p), ("scol_len", C.c_uint64)]


class HaloOut(C.Structure):
    """lstc_halo_out (include/lstc.h)."""
    _fields_ = [("off", C.c_void_p), ("col", C.c_void_p), ("l2g", C.c_void_p), ("n_local_cap", C.c_uint32),
                ("reserved", C.c_uint32), ("m_local_cap", C.c_uint64)]


class Request(C.Structure):
    """lstc_request (include/lstc.h)."""
    _fields_ = [("kind", C.c_uint32), ("scale", C.c_uint32), ("n", C.c_uint32), ("n_rows", C.c_uint32),
                ("m_raw", C.c_uint64), ("seed", C.c_uint64), ("a", C.c_double), ("b", C.c_double),
                ("c", C.c_double), ("off", C.c_void_p), ("col", C.c_void_p), ("m", C.c_uint64)]


class Result(C.Structure):
    """lstc_result (include/lstc.h)."""
    _fields_ = [("count", C.c_uint64), ("edges", C.c_uint64), ("latency_s", C.c_double)]


class WorkStats(C.Structure):
    """lstc_work (include/lstc.h)."""
    _fields_ = [(k, C.c_uint64) for k in ("rows", "items", "chunk_items", "tiny_rows", "hashed_keys", "segments",
                                          "tiny_segments", "streamed", "bytes")]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class ServerStats(C.Structure):
    """lstc_server_stats (include/lstc.h)."""
    _fields_ = [(k, C.c_uint64) for k in ("submitted", "completed", "rejected", "batches", "max_batch_graphs",
                                          "pool_bytes", "pool_used_bytes", "edges", "triangles")] + \
               [(k, C.c_double) for k in ("p50_s", "p99_s", "mean_s", "max_s")] + \
               [(k, C.c_uint64) for k in ("prep_groups", "max_prep_graphs")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


LSTC_REQ_RMAT, LSTC_REQ_ER, LSTC_REQ_CSR = 0, 1, 2

VP = C.c_void_p
U32, U64, I32, DBL, SZ = C.c_uint32, C.c_uint64, C.c_int, C.c_double, C.c_size_t
PU32, PU64, PSZ = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_size_t)
ST = C.c_int32

# name -> (restype, argtypes); every function declared in include/lstc.h
SIGNATURES = {
    "lstc_status_string": (C.c_char_p, [ST]),
    "lstc_last_error": (C.c_char_p, []),
    "lstc_version": (C.c_int, []),
    "lstc_selftest": (ST, [PU32]),
    "lstc_gen_rmat": (ST, [U32, U64, DBL, DBL, DBL, U64, VP, VP, VP]),
    "lstc_gen_er": (ST, [U32, U64, U64, VP, VP, VP]),
    "lstc_gen_chung_lu_workspace": (ST, [U32, PSZ]),
    "lstc_gen_chung_lu": (ST, [U32, DBL, DBL, DBL, U64, U64, VP, VP, VP, SZ, VP]),
    "lstc_gen_grid": (ST, [U32, U32, DBL, U64, VP, VP, VP]),
    "lstc_partition_random": (ST, [U32, U32, U64, VP, VP]),
    "lstc_partition_contig": (ST, [U32, U32, VP, VP]),
    "lstc_partition_read": (ST, [C.c_char_p, U32, U32, VP, VP]),
    "lstc_clean_orient_workspace": (ST, [U32, U64, PSZ]),
    "lstc_clean_orient": (ST, [VP, VP, U64, U32, I32, VP, VP, VP, VP, VP, VP, PU32, PU64, VP, SZ, VP]),
    "lstc_clean_orient_split": (ST, [VP, VP, U64, U32, I32, VP, VP, VP, VP, VP, VP, VP, VP, PU32, PU64, PU64, VP, SZ,
                                     VP]),
    "lstc_halo_workspace": (ST, [U32, PSZ]),
    "lstc_halo_sizes": (ST, [VP, VP, U32, VP, U32, I32, PU32, PU32, PU64, VP, SZ, VP]),
    "lstc_halo_extract": (ST, [VP, VP, U32, VP, U32, I32, VP, VP, VP, U32, U64, VP, SZ, VP]),
    "lstc_halo_parts_workspace": (ST, [U32, U32, PSZ]),
    "lstc_halo_parts_sizes": (ST, [VP, VP, U32, VP, U32, I32, PU32, PU32, PU64, VP, SZ, VP]),
    "lstc_halo_parts_extract": (ST, [VP, VP, U32, VP, U32, I32, C.POINTER(HaloOut), VP, SZ, VP]),
    "lstc_halo_parts_build": (ST, [VP, VP, U32, VP, U32, I32, C.POINTER(HaloOut), PU32, PU32, PU64, VP, SZ, VP]),
    "lstc_count_workspace": (ST, [U32, U32, U64, C.POINTER(Tune), PSZ]),
    "lstc_count": (ST, [VP, VP, U32, U32, U64, VP, VP, C.POINTER(Tune), VP, SZ, VP]),
    "lstc_count_split": (ST, [VP, VP, VP, VP, VP, VP, U32, U32, U64, VP, C.POINTER(Tune), VP, SZ, VP]),
    "lstc_count_split_stream": (ST, [VP, VP, VP, VP, VP, VP, VP, U64, U32, U32, U64, VP, C.POINTER(Tune), VP, SZ,
                                     VP]),
    "lstc_stream_copy_capacity": (U64, [U32, U64]),
    "lstc_stream_copy_workspace": (ST, [U32, PSZ]),
    "lstc_stream_copy": (ST, [VP, VP, U32, U64, U64, VP, VP, U64, VP, VP, U64, PU64, VP, SZ, VP]),
    "lstc_clean_orient_stream": (ST, [VP, VP, U64, U32, I32, VP, VP, VP, VP, VP, VP, VP, VP, VP, PU32, PU64, PU64,
                                      PU64, VP, SZ, VP]),
    "lstc_count_split_static_limit": (SZ, []),
    "lstc_count_split_static": (ST, [VP, VP, VP, VP, VP, VP, U32, U32, U64, VP, C.POINTER(Tune), VP]),
    "lstc_triangle_support": (ST, [VP, VP, U32, U32, U64, VP, VP, VP]),
    "lstc_in_lists_workspace": (ST, [U32, U32, U64, PSZ]),
    "lstc_in_lists": (ST, [VP, VP, U32, U32, U64, VP, VP, PU64, VP, SZ, VP]),
    "lstc_in_lists_segments": (ST, [VP, VP, U32, U32, U64, VP, VP, VP, VP, PU64, VP, SZ, VP]),
    "lstc_segments": (ST, [VP, VP, U32, U64, VP, U64, VP, VP, VP]),
    "lstc_count_batch_workspace": (ST, [C.POINTER(GraphDesc), U32, C.POINTER(Tune), PSZ]),
    "lstc_count_batch": (ST, [C.POINTER(GraphDesc), U32, VP, C.POINTER(Tune), VP, SZ, VP]),
    "lstc_batch_plan_create": (ST, [C.POINTER(GraphDesc), U32, C.POINTER(Tune), C.POINTER(VP)]),
    "lstc_batch_plan_run": (ST, [VP, VP, VP]),
    "lstc_batch_plan_launch_graph": (ST, [VP, VP, VP]),
    "lstc_batch_plan_destroy": (ST, [VP]),
    "lstc_shape_class": (ST, [C.POINTER(GraphDesc), U32, PU32]),
    "lstc_tune_resolve": (ST, [C.POINTER(GraphDesc), U32, C.POINTER(Tune), C.POINTER(Tune)]),
    "lstc_work_stats_workspace": (ST, [U32, PSZ]),
    "lstc_work_stats": (ST, [C.POINTER(GraphDesc), U32, C.POINTER(Tune), C.POINTER(WorkStats), VP, SZ, VP]),
    "lstc_timing_enable": (ST, [I32]),
    "lstc_timing_read": (ST, [C.POINTER(C.c_double), PU64]),
    "lstc_autotune": (ST, [C.POINTER(GraphDesc), U32, U32, C.POINTER(Tune), C.POINTER(C.c_double), VP]),
    "lstc_tune_cache_lookup": (ST, [C.c_char_p, U32, C.POINTER(Tune)]),
    "lstc_tune_cache_store": (ST, [C.c_char_p, U32, C.POINTER(Tune), DBL]),
    "lstc_server_create": (ST, [SZ, U32, C.POINTER(Tune), C.POINTER(VP)]),
    "lstc_server_submit": (ST, [VP, C.POINTER(Request), PU64]),
    "lstc_server_poll": (ST, [VP, U64, C.POINTER(Result)]),
    "lstc_server_wait": (ST, [VP, U64, DBL, C.POINTER(Result)]),
    "lstc_server_cancel": (ST, [VP, U64]),
    "lstc_triangle_support_split_workspace": (ST, [U32, U64, C.POINTER(C.c_size_t)]),
    "lstc_triangle_support_split": (ST, [VP, VP, VP, VP, U32, U32, U64, VP, VP, VP, C.c_size_t, VP]),
    "lstc_server_get_stats": (ST, [VP, C.POINTER(ServerStats)]),
    "lstc_server_destroy": (ST, [VP]),
}

_lib = None


def header_functions() -> list[str]:
    """Names of every function declared in include/lstc.h."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(lstc_[a-z0-9_]+)\s*\(", txt)))


def load(path: str | None = None):
    """Load liblstc.so (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing: build it with `python -m trianglecount_amd.build` "
                           "(the HIP library is required; there is no CPU fallback)")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(status: int, what: str = "") -> None:
    if status != LSTC_OK:
        lib = load()
        detail = lib.lstc_last_error().decode(errors="replace")
        raise LstcError(status, f"{what}: {lib.lstc_status_string(status).decode()} ({detail})")
