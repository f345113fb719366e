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


class LstcError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"lstc status {status}: {msg}")
        self.status = status


class Tune(C.Structure):
    """lstc_tune (include/lstc.h)."""
    _fields_ = [("wg_waves", C.c_uint32), ("table_scale", C.c_uint32),
                ("switch_ratio", C.c_uint32), ("split_segments", C.c_uint32),
                ("reserved", C.c_uint32 * 4)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_ if k != "reserved"}


class GraphDesc(C.Structure):
    """lstc_graph_desc (include/lstc.h)."""
    _fields_ = [("off", C.c_void_p), ("col", C.c_void_p), ("n_rows", C.c_uint32),
                ("n_owned", C.c_uint32)]


VP = C.c_void_p
U32, U64, I32, DBL, SZ = C.c_uint32, C.c_uint64, C.c_int, C.c_double, C.c_size_t
PU32, PU64, PSZ = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_size_t)
ST = C.c_int32

# name -> (restype, argtypes); every function declared in include/lstc.h
SIGNATURES = {
    "lstc_status_string": (C.c_char_p, [ST]),
    "lstc_last_error": (C.c_char_p, []),
    "lstc_version": (C.c_int, []),
    "lstc_gen_rmat": (ST, [U32, U64, DBL, DBL, DBL, U64, VP, VP, VP]),
    "lstc_gen_er": (ST, [U32, U64, U64, VP, VP, VP]),
    "lstc_gen_chung_lu_workspace": (ST, [U32, PSZ]),
    "lstc_gen_chung_lu": (ST, [U32, DBL, DBL, DBL, U64, U64, VP, VP, VP, SZ, VP]),
    "lstc_gen_grid": (ST, [U32, U32, DBL, U64, VP, VP, VP]),
    "lstc_partition_random": (ST, [U32, U32, U64, VP, VP]),
    "lstc_clean_orient_workspace": (ST, [U32, U64, PSZ]),
    "lstc_clean_orient": (ST, [VP, VP, U64, U32, I32, VP, VP, VP, VP, PU32, PU64, VP, SZ, VP]),
    "lstc_count_workspace": (ST, [U32, U32, U64, C.POINTER(Tune), PSZ]),
    "lstc_count": (ST, [VP, VP, U32, U32, U64, VP, VP, C.POINTER(Tune), VP, SZ, VP]),
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
