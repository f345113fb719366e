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
