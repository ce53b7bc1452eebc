"""C-ABI checks that need no GPU: the library loads, exports every symbol
include/sid_proj.h declares with the ctypes signatures the package uses,
and fails loudly (SID_ERR_NODEV) instead of computing anything on the CPU."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    src = open(os.path.join(ROOT, "include", "sid_proj.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(sid_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_bound_symbols():
    from sid_amd import _lib
    declared = _header_functions()
    assert declared, "no declarations parsed"
    assert sorted(_lib.SIGNATURES) == declared


def test_library_exports_every_declared_symbol():
    from sid_amd import _lib
    L = _lib.lib()
    for name in _header_functions():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    for name in _header_functions():
        assert re.search(r"\bT " + name + r"\b", out), name


def test_library_is_gfx950_code():
    from sid_amd import _lib
    blob = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_fails_loudly_without_device():
    import sid_amd
    if sid_amd.device_count() > 0:
        pytest.skip("a gfx950 device is visible")
    with pytest.raises(sid_amd.SidError, match="SID_ERR_NODEV"):
        sid_amd.PSF(np.ones((7, 7, 3, 3, 2), np.float32))
    with pytest.raises(sid_amd.SidError, match="SID_ERR_NODEV"):
        sid_amd.selftest()


def test_null_arguments_are_rejected():
    import ctypes
    from sid_amd import _lib
    L = _lib.lib()
    assert L.sid_psf_create(None, None, None, None) == 1   # SID_ERR_ARG
    assert L.sid_forward(None, None, None, None) == 1
    assert b"null" in L.sid_last_error()
    assert L.sid_version() >= 1
