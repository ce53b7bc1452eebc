"""Parity of the HIP projectors with the FP64 oracle (calls through the C ABI)."""
import zlib

import numpy as np
import pytest

from tests._helpers import c_backward, c_forward, c_isra, random_psf, rel_l2

pytestmark = pytest.mark.gpu

# FP32 tolerance: rel-L2 of one projection against FP64 (sums of <= 1e5
# positive FP32 products; observed ~1e-7).
TOL_FP32 = 2e-6


@pytest.fixture(scope="module")
def sid():
    import sid_amd
    if sid_amd.device_count() < 1:
        pytest.fail("no gfx950 device visible")
    return sid_amd


def test_selftest(sid):
    print("case: MFMA lane-layout selftest", flush=True)
    sid.selftest()


CASES = [
    # (n, X, Y, M, Nz, nz_fwd, crop)  -- X, Y not multiples of n exercise the last partial lenslet
    (3, 9, 12, 7, 2, 2, None),
    (3, 10, 8, 6, 2, 2, None),          # even M: forward offset M/2, backward M/2-1
    (5, 25, 30, 11, 3, 3, None),
    (5, 23, 27, 13, 3, 2, None),        # forward with fewer planes than H
    (3, 18, 15, 9, 3, 3, "crop"),       # forwardProjectACC crop semantics
    (15, 60, 75, 31, 2, 2, None),
    (15, 64, 47, 45, 2, 2, "crop"),
    (16, 48, 40, 21, 2, 2, None),       # largest lenslet (n*n = 256 columns, 16 column tiles)
    (2, 11, 9, 5, 2, 2, None),          # smallest column tile (n*n = 4 of 16 columns used)
    (3, 9, 9, 15, 2, 2, None),          # kernel wider than the sensor
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "n%d_X%d_Y%d_M%d_Nz%d_nz%d_%s" % (c[:6] + (c[6] or "full",)))
def test_forward_backward_vs_oracle(sid, case):
    n, X, Y, M, Nz, nz, crop = case
    print("case:", case, flush=True)
    rng = np.random.default_rng(zlib.crc32(repr(case).encode()))
    H = random_psf(rng, M, n, Nz)
    CA = None
    if crop:
        m = M // 2
        CA = np.array([[max(1, m + 1 - r), min(M, m + 1 + r)] for r in range(1, Nz + 1)], dtype=np.int32)
        CA[:, 1] = np.minimum(CA[:, 1], M)
    V = rng.random((X, Y, nz))
    O_ref = c_forward(H, V, CA)
    psf = sid.PSF(H.astype(np.float32), CA)
    O = psf.forward(V.astype(np.float32))
    assert O.shape == (X, Y)
    e = rel_l2(O, O_ref)
    print("forward rel-L2", e, flush=True)
    assert e < TOL_FP32
    W = rng.random((X, Y))
    B_ref = c_backward(H, W, CA)
    B = psf.backward(W.astype(np.float32))
    assert B.shape == (X, Y, Nz)
    e = rel_l2(B, B_ref)
    print("backward rel-L2", e, flush=True)
    assert e < TOL_FP32
    psf.close()


def test_adjoint_identity(sid):
    """<F v, w> = <v, B w> on the device results."""
    print("case: adjoint n15 X90 Y105 M61 Nz3", flush=True)
    rng = np.random.default_rng(7)
    H = random_psf(rng, 61, 15, 3).astype(np.float32)
    psf = sid.PSF(H)
    v = rng.random((90, 105, 3)).astype(np.float32)
    w = rng.random((90, 105)).astype(np.float32)
    lhs = float(np.sum(psf.forward(v).astype(np.float64) * w))
    rhs = float(np.sum(v.astype(np.float64) * psf.backward(w)))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


def test_isra_vs_oracle(sid):
    print("case: isra n5 X40 Y35 M15 Nz3 8 iters", flush=True)
    rng = np.random.default_rng(3)
    H = random_psf(rng, 15, 5, 3)
    V = rng.random((40, 35, 3))
    y = c_forward(H, V)
    htf = c_backward(H, y)
    X_ref = c_isra(H, htf, htf, 8)
    psf = sid.PSF(H.astype(np.float32))
    X = psf.deconv_rl(htf.astype(np.float32), 8, htf.astype(np.float32))
    e = rel_l2(X, X_ref)
    print("isra rel-L2", e, flush=True)
    assert e < 1e-5


def test_errors(sid):
    rng = np.random.default_rng(0)
    H = random_psf(rng, 7, 3, 2).astype(np.float32)
    psf = sid.PSF(H)
    with pytest.raises(sid.SidError, match="SID_ERR_NZ"):
        psf.forward(np.zeros((9, 9, 3), np.float32))
    with pytest.raises(ValueError):
        sid.PSF(np.zeros((7, 7, 3), np.float32))


def test_fp64_mode_vs_oracle(sid):
    """The FP64 device path agrees with the C oracle to FP64 round-off (float I/O)."""
    print("case: fp64 mode n5 X40 Y35 M15 Nz3", flush=True)
    rng = np.random.default_rng(11)
    H = random_psf(rng, 15, 5, 3).astype(np.float32).astype(np.float64)
    V = rng.random((40, 35, 3)).astype(np.float32).astype(np.float64)
    psf = sid.PSF(H.astype(np.float32), precision="fp64")
    e = rel_l2(psf.forward(V.astype(np.float32)), c_forward(H, V))
    assert e < 1e-7, e   # bounded by the float output rounding
    y = c_forward(H, V)
    htf = c_backward(H, y).astype(np.float32).astype(np.float64)
    X = psf.deconv_rl(htf.astype(np.float32), 6, htf.astype(np.float32))
    e = rel_l2(X, c_isra(H, htf, htf, 6))
    print("fp64 isra rel-L2", e, flush=True)
    assert e < 1e-6, e


@pytest.mark.parametrize("prec,tol_fwd", [("bf16x2", 3e-3), ("bf16x3", 2e-5)])
def test_bf16_split_error_budget(sid, prec, tol_fwd):
    """BF16 split products: H rounded to BF16 (2^-9 relative) dominates bf16x2;
    bf16x3 adds the H lo part and is near FP32."""
    print("case:", prec, "n15 X90 Y75 M61 Nz3", flush=True)
    from sid_amd import synth
    H, _ = synth.make_psf(15, 61, (-3, 0, 3), dtype=np.float64)
    rng = np.random.default_rng(5)
    V = rng.random((90, 75, 3))
    O_ref = c_forward(H, V)
    psf = sid.PSF(H.astype(np.float32), precision=prec)
    e = rel_l2(psf.forward(V.astype(np.float32)), O_ref)
    print(prec, "forward rel-L2", e, flush=True)
    assert e < tol_fwd
    B_ref = c_backward(H, O_ref)
    e = rel_l2(psf.backward(O_ref.astype(np.float32)), B_ref)
    print(prec, "backward rel-L2", e, flush=True)
    assert e < tol_fwd
