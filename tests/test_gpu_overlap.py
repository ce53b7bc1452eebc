"""The two halves of a projection and the split between them.

A mixed plan runs the direct GEMM on the handle's second stream beside the FFT
path on the caller's stream (fork / join).  The halves write disjoint buffers
and the reduction sums them in a fixed order, so the results must be
bit-identical to running the halves one after the other (SID_NO_OVERLAP), and a
GEMM grid capped below the item count (blocks looping over items,
SID_GEMM_BLOCKS) must give bit-identical results to one block per item.
Pinned mixed plans agree with the FP64 oracle; invalid pins are rejected."""
import ctypes

import numpy as np
import pytest

from tests._helpers import c_backward, c_forward, c_isra, random_psf, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    import sid_amd
    if sid_amd.device_count() < 1:
        pytest.fail("no gfx950 device visible")
    return torch, sid_amd


def _run(sid_amd, H, V, W, htf, precision, pins):
    psf = sid_amd.PSF(H.astype(np.float32), precision=precision)
    psf.set_plane_methods(pins)
    out = (psf.forward(V.astype(np.float32)), psf.backward(W.astype(np.float32)),
           psf.deconv_rl(htf.astype(np.float32), 3, htf.astype(np.float32)))
    from sid_amd._lib import check, lib
    stats = (ctypes.c_int64 * 8)()
    check(lib().sid_plan_stats(psf._h, stats))
    meth = psf.info()["plane_method"]
    psf.close()
    return out, list(stats), meth


CASES = [
    # (n, X, Y, M, Nz, pins)
    (5, 33, 41, 13, 4, ("direct", "fft", "direct", "fft")),
    (15, 90, 105, 31, 3, ("fft", "direct", "direct")),
]


@pytest.mark.parametrize("precision", ["fp32", "bf16x2", "bf16x3"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "n%d_X%d_Y%d_M%d_Nz%d" % c[:5])
def test_overlapped_halves_bit_identical(env, case, precision, monkeypatch):
    torch, sid_amd = env
    n, X, Y, M, Nz, pins = case
    print("case: overlap", precision, case, flush=True)
    rng = np.random.default_rng(5 + n)
    H = random_psf(rng, M, n, Nz)
    V, W = rng.random((X, Y, Nz)), rng.random((X, Y))
    htf = c_backward(H, c_forward(H, V))
    runs = {}
    for mode, envs in (("serial", {"SID_NO_OVERLAP": "1"}), ("overlap", {}), ("capped", {"SID_GEMM_BLOCKS": "3"})):
        for k in ("SID_NO_OVERLAP", "SID_GEMM_BLOCKS"):
            monkeypatch.delenv(k, raising=False)
        for k, v in envs.items():
            monkeypatch.setenv(k, v)
        runs[mode] = _run(sid_amd, H, V, W, htf, precision, list(pins))
    for k in ("SID_NO_OVERLAP", "SID_GEMM_BLOCKS"):
        monkeypatch.delenv(k, raising=False)
    out_s, stats, meth = runs["serial"]
    assert meth == list(pins)
    assert stats[2] > 3 and stats[3] > 3, "the capped grid must be smaller than the item count"
    for mode in ("overlap", "capped"):
        for a, b in zip(runs[mode][0], out_s):
            assert np.array_equal(a, b), mode
    if precision == "fp32":
        tol = 1e-5
        assert rel_l2(out_s[0], c_forward(H, V)) < tol
        assert rel_l2(out_s[1], c_backward(H, W)) < tol
        assert rel_l2(out_s[2], c_isra(H, htf, htf, 3)) < 2 * tol


def test_plane_pins(env):
    torch, sid_amd = env
    from sid_amd._lib import SidError
    print("case: plane pins", flush=True)
    rng = np.random.default_rng(3)
    n, X, Y, M, Nz = 3, 12, 15, 7, 3
    H = random_psf(rng, M, n, Nz)
    V = rng.random((X, Y, Nz))
    psf = sid_amd.PSF(H.astype(np.float32))
    ref = c_forward(H, V)
    for pins in (["fft", "fft", "direct"], ["direct", "auto", "fft"], ["direct"] * 3):
        psf.set_plane_methods(pins)
        out = psf.forward(V.astype(np.float32))
        meth = psf.info()["plane_method"]
        for c, m in enumerate(pins):
            if m != "auto":
                assert meth[c] == m
        assert rel_l2(out, ref) < 1e-5
    with pytest.raises(SidError):
        psf.set_plane_methods(["fft", "fft"])          # one entry per plane
    psf.set_plane_methods(None)
    psf.set_method("direct")                           # a forced method wins over pins
    psf.set_plane_methods(["fft"] * 3)
    psf.forward(V.astype(np.float32))
    assert psf.info()["plane_method"] == ["direct"] * 3
    psf.close()
    # even supports cannot take the FFT path
    He = random_psf(rng, 6, n, 2)
    pe = sid_amd.PSF(He.astype(np.float32))
    with pytest.raises(SidError):
        pe.set_plane_methods(["fft", "direct"])
    pe.close()
