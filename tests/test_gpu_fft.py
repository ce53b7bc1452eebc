"""FFT path: the hand-written Stockham FFT against numpy, and the FFT
projector (sublattice-periodised spectra) against the FP64 oracle."""
import ctypes

import zlib

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


@pytest.mark.parametrize("L", [2, 3, 4, 5, 15, 30, 48, 80, 960, 1200, 2048])
def test_fft_primitive_vs_numpy(env, L):
    torch, sid_amd = env
    from sid_amd._lib import check, lib
    rng = np.random.default_rng(L)
    nseq = 7
    x = (rng.standard_normal((nseq, L)) + 1j * rng.standard_normal((nseq, L))).astype(np.complex64)
    for direction, ref in ((-1, np.fft.fft(x, axis=1)), (1, np.fft.ifft(x, axis=1) * L)):
        t = torch.from_numpy(np.ascontiguousarray(x).view(np.float32)).cuda()
        torch.cuda.synchronize()
        check(lib().sid_fft_device(L, nseq, ctypes.c_void_p(t.data_ptr()), direction, None))
        got = t.cpu().numpy().view(np.complex64)
        assert rel_l2(got, ref) < 2e-6 * max(1.0, np.log2(L))


CASES = [
    # (n, X, Y, M, Nz, crop)
    (3, 9, 12, 7, 2, None),
    (5, 25, 30, 11, 3, None),
    (5, 23, 27, 13, 3, None),           # X, Y not multiples of n
    (3, 18, 15, 9, 3, "crop"),          # odd crops
    (15, 60, 75, 31, 2, None),
    (15, 90, 105, 61, 3, None),
    (3, 10, 8, 6, 2, None),             # even M: planes stay direct
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "n%d_X%d_Y%d_M%d_Nz%d_%s" % (c[:5] + (c[5] or "full",)))
def test_fft_projector_vs_oracle(env, case):
    torch, sid_amd = env
    n, X, Y, M, Nz, crop = case
    print("case: fft", case, flush=True)
    rng = np.random.default_rng(zlib.crc32(repr(case).encode()))
    H = random_psf(rng, M, n, Nz)
    CA = None
    if crop:
        m = M // 2
        CA = np.array([[max(1, m + 1 - r), min(M, m + 1 + r)] for r in range(1, Nz + 1)], dtype=np.int32)
    V = rng.random((X, Y, Nz))
    psf = sid_amd.PSF(H.astype(np.float32), CA, method="fft")
    e_f = rel_l2(psf.forward(V.astype(np.float32)), c_forward(H, V, CA))
    W = rng.random((X, Y))
    e_b = rel_l2(psf.backward(W.astype(np.float32)), c_backward(H, W, CA))
    y = c_forward(H, V, CA)
    htf = c_backward(H, y, CA)
    X4 = psf.deconv_rl(htf.astype(np.float32), 4, htf.astype(np.float32))
    e_r = rel_l2(X4, c_isra(H, htf, htf, 4, CA))
    print("fft rel-L2 forward", e_f, "backward", e_b, "isra", e_r, flush=True)
    assert e_f < 1e-5 and e_b < 1e-5 and e_r < 2e-5
    psf.close()


def test_fft_and_direct_agree_on_workload(env):
    """Synthetic light-field PSF (n = 15, 3 planes, M = 105): FFT and direct
    projections agree, BF16x2 direct planes with FFT planes stay within 1e-3."""
    torch, sid_amd = env
    from sid_amd import synth
    print("case: fft vs direct, synthetic PSF n15 M105", flush=True)
    H, _ = synth.make_psf(15, 105, (-6, 0, 6))
    rng = np.random.default_rng(9)
    V = rng.random((120, 135, 3)).astype(np.float32)
    pd = sid_amd.PSF(H, method="direct")
    pf = sid_amd.PSF(H, method="fft")
    Od, Of = pd.forward(V), pf.forward(V)
    assert rel_l2(Of, Od) < 1e-5
    assert rel_l2(pf.backward(Od), pd.backward(Od)) < 1e-5


@pytest.mark.parametrize("case", [CASES[1], CASES[3], CASES[5]],
                         ids=lambda c: "n%d_X%d_Y%d_M%d_Nz%d_%s" % (c[:5] + (c[5] or "full",)))
def test_fft_fp16_spectra_vs_oracle(env, case, monkeypatch):
    """BF16X2 mode with spectra of the whole padded sensor (SID_FFT_UNTILED;
    the tiled form has its own tests) keeps them in FP16, row q scaled by
    1/sum|h_q|.  One stored value carries a relative rounding error of at most
    2^-11 (4.9e-4); the bar is that bound, and every output sums 225 * Nz
    such values so the measured error sits well below it."""
    torch, sid_amd = env
    monkeypatch.setenv("SID_FFT_UNTILED", "1")
    n, X, Y, M, Nz, crop = case
    print("case: fft fp16 spectra", case, flush=True)
    rng = np.random.default_rng(zlib.crc32(repr(case).encode()) + 1)
    H = random_psf(rng, M, n, Nz)
    CA = None
    if crop:
        m = M // 2
        CA = np.array([[max(1, m + 1 - r), min(M, m + 1 + r)] for r in range(1, Nz + 1)], dtype=np.int32)
    V = rng.random((X, Y, Nz))
    psf = sid_amd.PSF(H.astype(np.float32), CA, precision="bf16x2", method="fft")
    e_f = rel_l2(psf.forward(V.astype(np.float32)), c_forward(H, V, CA))
    b16 = psf.info()["spectra_bytes"]
    assert b16 > 0
    W = rng.random((X, Y))
    e_b = rel_l2(psf.backward(W.astype(np.float32)), c_backward(H, W, CA))
    print("fp16-spectra rel-L2 forward", e_f, "backward", e_b, flush=True)
    assert e_f < 2.0 ** -11 and e_b < 2.0 ** -11
    # switching to FP32 rebuilds FP32 spectra (twice the bytes) and tightens the result
    psf.set_precision("fp32")
    e32 = rel_l2(psf.forward(V.astype(np.float32)), c_forward(H, V, CA))
    assert psf.info()["spectra_bytes"] == 2 * b16
    assert e32 < 1e-5
    psf.close()


@pytest.mark.parametrize("precision", ["fp32", "bf16x2"])
def test_fft_fused_and_unfused_sub_transforms_agree(env, precision, monkeypatch):
    """The LDS-resident pair transform (two real sub-images as one complex
    2-D DFT) and the separate gather / 1-D FFT / pack kernels compute the same
    projections; both agree with the oracle."""
    torch, sid_amd = env
    n, X, Y, M, Nz = 5, 33, 41, 13, 3        # 15 sub-images: an unpaired last one
    print("case: fused vs unfused", precision, flush=True)
    rng = np.random.default_rng(77)
    H = random_psf(rng, M, n, Nz)
    V = rng.random((X, Y, Nz))
    W = rng.random((X, Y))
    out = {}
    for mode in ("fused", "unfused"):
        if mode == "unfused":
            monkeypatch.setenv("SID_FFT_UNFUSED", "1")
        psf = sid_amd.PSF(H.astype(np.float32), precision=precision, method="fft")
        out[mode] = (psf.forward(V.astype(np.float32)), psf.backward(W.astype(np.float32)))
        psf.close()
    monkeypatch.delenv("SID_FFT_UNFUSED", raising=False)
    tol = 1e-5 if precision == "fp32" else 2.0 ** -11
    for i, ref in enumerate((c_forward(H, V), c_backward(H, W))):
        e_fu, e_un = rel_l2(out["fused"][i], ref), rel_l2(out["unfused"][i], ref)
        print(precision, "fused", e_fu, "unfused", e_un, flush=True)
        assert e_fu < tol and e_un < tol
        assert rel_l2(out["fused"][i], out["unfused"][i]) < 1e-5
