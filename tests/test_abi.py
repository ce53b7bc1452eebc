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


def test_patched_rect_matches_oracle_project_forward_patched():
    """The library's rectangle rule is integer-identical to
    project_forward_patched (oracle) for random boxes, including clipped ones."""
    from oracle import sid_oracle as O
    from sid_amd import library
    rng = np.random.default_rng(0)
    for _ in range(300):
        n = int(rng.choice([3, 5, 15]))
        M = int(rng.choice([7, 9, 21, 45]))
        X, Y = int(rng.integers(20, 120)), int(rng.integers(20, 120))
        x0 = int(rng.integers(1, X + 1)); x1 = int(rng.integers(x0, min(X, x0 + 25) + 1))
        y0 = int(rng.integers(1, Y + 1)); y1 = int(rng.integers(y0, min(Y, y0 + 25) + 1))
        box = (x0, x1, y0, y1, 1, 1)
        rect = library.patched_rect(n, M, (X, Y), box)
        # oracle: the rectangle that project_forward_patched pastes
        pixels = np.array([[x0, y0, 1], [x1, y1, 1]])
        H = np.zeros((M, M, n, n, 1))
        _, info = O.project_forward_patched(pixels, (X, Y), n, H, np.ones((1, 1, 1)),
                                            np.ones((2, 3), dtype=np.int64))
        assert rect == (info["x0"], info["x1"], info["y0"], info["y1"]), (n, M, X, Y, box)


def test_library_placement_matches_oracle():
    from oracle import sid_oracle as O
    from sid_amd import library
    rng = np.random.default_rng(1)
    for _ in range(300):
        c = (rng.uniform(-1, 60), rng.uniform(-1, 60), rng.uniform(0, 12))
        r = int(rng.integers(1, 8))
        axial = float(rng.choice([1.0, 2.0, 2.5, 4.0]))
        p = O.library_gpu_placement(c, r, axial, (50, 55), 10)
        dz = int(np.ceil(r / axial))
        assert library.placement(c, r, dz, (50, 55), 10) == (p["x0"], p["x1"], p["y0"], p["y1"], p["z0"], p["z1"])
    # half-integers round away from zero
    assert library.placement((2.5, -0.5, 3.5), 1, 0, (10, 10), 9)[0:2] == (2, 4)
