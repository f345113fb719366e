"""CPU checks of the C ABI: the library loads and exports every declared symbol.

No compute entry point is called here (there is no GPU in the build container).
"""
import ctypes
import os
import subprocess

import pytest

from trianglecount_amd import _lib


def test_library_built_and_loads():
    assert os.path.exists(_lib.LIB_PATH), "run python -m trianglecount_amd.build"
    _lib.load()


def test_every_header_function_is_exported():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _lib.header_functions()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_header():
    assert set(_lib.header_functions()) == set(_lib.SIGNATURES)


def test_exports_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    for n in _lib.header_functions():
        assert n in exported, n


def test_status_strings_host_only():
    lib = _lib.load()
    assert lib.lstc_status_string(0) == b"ok"
    assert lib.lstc_status_string(-5).startswith(b"workspace")
    assert lib.lstc_status_string(12345) == b"unknown status"
    assert lib.lstc_version() >= 10000


def test_check_raises_with_detail():
    with pytest.raises(_lib.LstcError) as e:
        _lib.check(_lib.LSTC_ERR_INVALID, "probe")
    assert e.value.status == _lib.LSTC_ERR_INVALID


def test_workspace_queries_are_host_only():
    lib = _lib.load()
    nb = ctypes.c_size_t()
    assert lib.lstc_count_workspace(1000, 1000, 5000, None, ctypes.byref(nb)) == 0
    assert nb.value > 0
    assert lib.lstc_clean_orient_workspace(1000, 5000, ctypes.byref(nb)) == 0
    assert nb.value >= 2 * 5000 * 8 * 2
    assert lib.lstc_count_workspace(1, 1, 0, None, None) == _lib.LSTC_ERR_INVALID


# Shape validation happens on the host before any HIP call, so these run without
# a device.  The pointers are placeholders that are never dereferenced.
_FAKE = ctypes.c_void_p(0x10000)


def _status(st):
    return _lib.load().lstc_status_string(st).decode()


def test_count_rejects_bad_shapes_on_host():
    lib = _lib.load()
    ws = ctypes.c_size_t(1 << 20)
    # n_owned > n_rows
    assert lib.lstc_count(_FAKE, _FAKE, 10, 11, 20, _FAKE, None, None, _FAKE, ws, None) == _lib.LSTC_ERR_INVALID
    # m beyond the 32-bit work positions
    assert lib.lstc_count(_FAKE, _FAKE, 10, 10, 0xFFFFF000, _FAKE, None, None, _FAKE, ws, None) == _lib.LSTC_ERR_INVALID
    # no output
    assert lib.lstc_count(_FAKE, _FAKE, 10, 10, 20, None, None, None, _FAKE, ws, None) == _lib.LSTC_ERR_INVALID
    # split count without in-lists
    assert lib.lstc_count_split(_FAKE, _FAKE, None, None, None, None, 10, 10, 20, _FAKE, None, _FAKE, ws,
                                None) == _lib.LSTC_ERR_INVALID
    assert "lstc_count" in lib.lstc_last_error().decode()
    # segment words: both or neither, and only with in-lists
    assert lib.lstc_count_split(_FAKE, _FAKE, _FAKE, _FAKE, _FAKE, None, 10, 10, 20, _FAKE, None, _FAKE, ws,
                                None) == _lib.LSTC_ERR_INVALID
    assert "segment" in lib.lstc_last_error().decode()
    # segments: more in-list entries than arcs, NULL arrays
    assert lib.lstc_segments(_FAKE, _FAKE, 10, 20, _FAKE, 21, _FAKE, _FAKE, None) == _lib.LSTC_ERR_INVALID
    assert lib.lstc_segments(_FAKE, _FAKE, 10, 20, None, 5, _FAKE, _FAKE, None) == _lib.LSTC_ERR_INVALID
    assert lib.lstc_segments(_FAKE, _FAKE, 10, 0xFFFFF000, None, 0, _FAKE, None, None) == _lib.LSTC_ERR_INVALID


def test_stream_copy_entries_reject_bad_shapes_on_host():
    """lstc_clean_orient_stream / lstc_count_split_stream refuse, before any device call, a missing
    copy, a copy without segment words and sizes beyond their 32-bit positions."""
    lib = _lib.load()
    ws = ctypes.c_size_t(1 << 20)
    assert lib.lstc_stream_copy_capacity(10, 100) == 100 + 31 * 10
    u32, u64, u64b, u64c = ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    args = lambda scol, m_raw, n: (_FAKE, _FAKE, m_raw, n, 1, None, None, _FAKE, _FAKE, _FAKE, _FAKE, _FAKE, _FAKE,
                                   scol, ctypes.byref(u32), ctypes.byref(u64), ctypes.byref(u64b),
                                   ctypes.byref(u64c), _FAKE, ws, None)
    assert lib.lstc_clean_orient_stream(*args(None, 100, 10)) == _lib.LSTC_ERR_INVALID
    assert "streaming copy" in lib.lstc_last_error().decode()
    assert lib.lstc_clean_orient_stream(*args(_FAKE, 1 << 32, 1 << 27)) == _lib.LSTC_ERR_INVALID
    assert "2^32" in lib.lstc_last_error().decode()
    cs = lambda scol, seg, scol_len: (_FAKE, _FAKE, _FAKE, _FAKE, seg, seg, scol, scol_len, 10, 10, 20, _FAKE, None,
                                      _FAKE, ws, None)
    assert lib.lstc_count_split_stream(*cs(None, _FAKE, 40)) == _lib.LSTC_ERR_INVALID
    assert lib.lstc_count_split_stream(*cs(_FAKE, None, 40)) == _lib.LSTC_ERR_INVALID
    assert lib.lstc_count_split_stream(*cs(_FAKE, _FAKE, 0xFFFFF000)) == _lib.LSTC_ERR_INVALID
    assert "streaming copy" in lib.lstc_last_error().decode()


def test_stream_copy_of_a_csr_rejects_bad_shapes_on_host():
    lib = _lib.load()
    ws = ctypes.c_size_t(1 << 20)
    used = ctypes.c_uint64()
    sc = lambda m, n_rows, words, cap: (_FAKE, _FAKE, n_rows, m, words, _FAKE, _FAKE, 4, _FAKE, _FAKE, cap,
                                        ctypes.byref(used), _FAKE, ws, None)
    assert lib.lstc_stream_copy(*sc(100, 10, 100, 409)) == _lib.LSTC_ERR_INVALID  # capacity < m + 31 n_rows
    assert "capacity" in lib.lstc_last_error().decode()
    assert lib.lstc_stream_copy(*sc(1 << 32, 10, 0, 1 << 33)) == _lib.LSTC_ERR_INVALID
    assert lib.lstc_stream_copy(*sc(100, 10, 101, 410)) == _lib.LSTC_ERR_INVALID  # more words than arcs
    assert lib.lstc_stream_copy(_FAKE, None, 10, 100, 0, None, None, 0, None, _FAKE, 410, ctypes.byref(used),
                                _FAKE, ws, None) == _lib.LSTC_ERR_INVALID
    assert "NULL" in lib.lstc_last_error().decode()


def test_batch_and_in_lists_reject_bad_shapes_on_host():
    lib = _lib.load()
    descs = (_lib.GraphDesc * 2)()
    for d in descs:
        d.off, d.col, d.m, d.n_rows, d.n_owned = 0x10000, 0x10000, 20, 10, 10
    descs[1].in_off = 0x10000  # in_off without in_col
    nb = ctypes.c_size_t()
    assert lib.lstc_count_batch(descs, 2, _FAKE, None, _FAKE, ctypes.c_size_t(1 << 20), None) == _lib.LSTC_ERR_INVALID
    descs[1].in_col, descs[1].col_seg = 0x10000, 0x10000  # one segment-word array without the other
    assert lib.lstc_count_batch(descs, 2, _FAKE, None, _FAKE, ctypes.c_size_t(1 << 20), None) == _lib.LSTC_ERR_INVALID
    assert "segment" in lib.lstc_last_error().decode()
    big = (_lib.GraphDesc * 2048)()
    assert lib.lstc_count_batch(big, 2048, _FAKE, None, _FAKE, ctypes.c_size_t(1 << 20), None) == _lib.LSTC_ERR_INVALID
    assert lib.lstc_in_lists_workspace(10, 11, 20, ctypes.byref(nb)) == _lib.LSTC_ERR_INVALID
    k = ctypes.c_uint64()
    assert lib.lstc_in_lists(_FAKE, _FAKE, 10, 11, 20, _FAKE, _FAKE, ctypes.byref(k), _FAKE, ctypes.c_size_t(1 << 20),
                             None) == _lib.LSTC_ERR_INVALID
    assert lib.lstc_in_lists_workspace(10, 10, 20, ctypes.byref(nb)) == _lib.LSTC_OK and nb.value > 0
    # the fused build: the same checks, plus the segment words' 32-bit starts and both arrays
    big_ws = ctypes.c_size_t(1 << 20)
    assert lib.lstc_in_lists_segments(_FAKE, _FAKE, 10, 11, 20, _FAKE, _FAKE, _FAKE, _FAKE, ctypes.byref(k), _FAKE,
                                      big_ws, None) == _lib.LSTC_ERR_INVALID
    assert lib.lstc_in_lists_segments(_FAKE, _FAKE, 10, 10, 0xFFFFF000, _FAKE, _FAKE, _FAKE, _FAKE, ctypes.byref(k),
                                      _FAKE, big_ws, None) == _lib.LSTC_ERR_INVALID
    assert lib.lstc_in_lists_segments(_FAKE, _FAKE, 10, 10, 20, _FAKE, _FAKE, _FAKE, None, ctypes.byref(k), _FAKE,
                                      big_ws, None) == _lib.LSTC_ERR_INVALID
    assert "segment" in lib.lstc_last_error().decode()
    assert lib.lstc_in_lists_segments(_FAKE, _FAKE, 10, 10, 20, _FAKE, _FAKE, _FAKE, _FAKE, ctypes.byref(k), _FAKE,
                                      ctypes.c_size_t(16), None) == _lib.LSTC_ERR_WORKSPACE


def test_support_rejects_bad_shapes_on_host():
    lib = _lib.load()
    ws = ctypes.c_size_t(1 << 30)
    for args in ((10, 11, 20, _FAKE, _FAKE),             # n_owned > n_rows
                 (10, 10, 20, None, None),               # no output
                 (10, 10, 0xFFFFF000, _FAKE, None)):     # m beyond the walk's 32-bit positions
        n, no, m, sup, pv = args
        assert lib.lstc_triangle_support(_FAKE, _FAKE, n, no, m, sup, pv, None) == _lib.LSTC_ERR_INVALID, args
        assert lib.lstc_triangle_support_split(_FAKE, _FAKE, _FAKE, _FAKE, n, no, m, sup, pv, _FAKE, ws,
                                               None) == _lib.LSTC_ERR_INVALID, args
    # the split walk needs the in-lists and its workspace
    assert lib.lstc_triangle_support_split(_FAKE, _FAKE, None, None, 10, 10, 20, _FAKE, None, _FAKE, ws,
                                           None) == _lib.LSTC_ERR_INVALID
    nb = ctypes.c_size_t()
    assert lib.lstc_triangle_support_split_workspace(10, 20, ctypes.byref(nb)) == _lib.LSTC_OK and nb.value > 0
    assert lib.lstc_triangle_support_split(_FAKE, _FAKE, _FAKE, _FAKE, 10, 10, 20, _FAKE, None, _FAKE,
                                           ctypes.c_size_t(nb.value - 1), None) == _lib.LSTC_ERR_WORKSPACE


def _descs(sizes):
    arr = (_lib.GraphDesc * len(sizes))()
    for i, (n, m) in enumerate(sizes):
        arr[i].n_rows = arr[i].n_owned = n
        arr[i].m = m
    return arr


def _resolve(lib, descs, tune=None):
    out = _lib.Tune()
    t = ctypes.byref(tune) if tune is not None else None
    assert lib.lstc_tune_resolve(descs, len(descs), t, ctypes.byref(out)) == 0
    return out.as_dict()


def test_tune_resolve_host_only():
    """The tuning a launch runs with: defaults, clamps and the small-launch cap of small_max."""
    lib = _lib.load()
    small = _descs([(1000, 1 << 20), (5000, 1 << 21)])  # 3 * 2^20 arcs < 2^23
    big = _descs([(1 << 20, 1 << 23)])
    d = _resolve(lib, small)
    assert d == {"table_scale": 6, "small_max": 128, "small_group": 0, "grid_per_cu": 8, "unroll": 4,
                 "tiny_max": 8, "waves": 4}
    # the workgroup size: 2, 4 or 8 waves; the default grid keeps 32 waves per CU
    for w, want_w, want_grid in ((2, 2, 16), (8, 8, 4), (3, 2, 16), (16, 8, 4)):
        r = _resolve(lib, big, _lib.Tune(waves=w))
        assert (r["waves"], r["grid_per_cu"]) == (want_w, want_grid), w
    assert _resolve(lib, big)["small_max"] == 170  # min(256, 1024 // 6)
    assert _resolve(lib, big, _lib.Tune(table_scale=2))["small_max"] == 256
    assert _resolve(lib, big, _lib.Tune(table_scale=16))["small_max"] == 64
    # an explicit small_max is kept (within its clamp) even for a small launch
    assert _resolve(lib, small, _lib.Tune(small_max=256))["small_max"] == 170
    assert _resolve(lib, small, _lib.Tune(small_max=64))["small_max"] == 64
    # clamps
    c = _resolve(lib, big, _lib.Tune(table_scale=40, grid_per_cu=99, unroll=5, tiny_max=1, small_group=1))
    assert (c["table_scale"], c["grid_per_cu"], c["unroll"], c["tiny_max"], c["small_group"]) == (16, 32, 4, 1, 4)
    # a resolved tune resolves to itself: replaying it runs what was resolved
    for descs in (small, big):
        for t in (None, _lib.Tune(table_scale=4), _lib.Tune(small_max=100, unroll=8, tiny_max=4),
                  _lib.Tune(waves=8, grid_per_cu=3)):
            r = _resolve(lib, descs, t)
            assert _resolve(lib, descs, _lib.Tune(**r)) == r
    assert lib.lstc_tune_resolve(small, 2, None, None) == _lib.LSTC_ERR_INVALID


def test_static_count_refuses_before_any_device_call():
    """lstc_count_split_static checks its arguments and the scratch limit on the host first."""
    lib = _lib.load()
    assert lib.lstc_count_split_static_limit() == 32 << 20
    fake = ctypes.c_void_p(0x1000)
    # scratch need above the limit
    st = lib.lstc_count_split_static(fake, fake, fake, fake, None, None, 1 << 23, 1 << 23, 1 << 20, fake, None, None)
    assert st == _lib.LSTC_ERR_INVALID and b"static scratch" in lib.lstc_last_error()
    # missing in-lists / output
    assert lib.lstc_count_split_static(fake, fake, None, None, None, None, 10, 10, 20, fake, None, None) == \
        _lib.LSTC_ERR_INVALID
    assert lib.lstc_count_split_static(fake, fake, fake, fake, None, None, 10, 10, 20, None, None, None) == \
        _lib.LSTC_ERR_INVALID
