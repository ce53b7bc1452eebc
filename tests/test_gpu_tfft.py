"""The tiled FFT path (BF16X2 precision): overlap-save tiles of T x T lenslets,
per-bin complex GEMMs on MFMA with FP16 spectra and FP16 hi + lo inputs,
against the FP64 oracle.  Covers partial lenslets, cropped planes with
different tap ranges, Nnum from 2 to 16, every instantiated tile length, one
to four 16-tile column blocks of the GEMM, a forward input with fewer planes
than H, mixed direct + FFT plans and the ISRA loop."""
import zlib

import numpy as np
import pytest

from tests._helpers import c_backward, c_forward, c_isra, random_psf, rel_l2

pytestmark = pytest.mark.gpu

# One stored FP16 spectrum value carries a relative rounding error of at most
# 2^-11; the inputs' hi + lo split keeps 21 bits.  Every output sums many such
# products, so the measured error sits well below this bar.
TOL = 2.0 ** -11


@pytest.fixture(scope="module")
def env():
    import torch
    import sid_amd
    if sid_amd.device_count() < 1:
        pytest.fail("no gfx950 device visible")
    return torch, sid_amd


def _crop(M, Nz):
    m = M // 2
    return np.array([[max(1, m + 1 - r), min(M, m + 1 + r)] for r in range(1, Nz + 1)], dtype=np.int32)


CASES = [
    # (n, X, Y, M, Nz, crop)
    (3, 9, 12, 7, 2, None),
    (5, 25, 30, 11, 3, None),
    (5, 23, 27, 13, 3, None),            # X, Y not multiples of n
    (3, 18, 15, 9, 3, "crop"),           # odd crops: planes with different tap ranges
    (15, 60, 75, 31, 2, None),
    (15, 90, 105, 61, 3, None),
    (4, 44, 36, 13, 3, None),            # even Nnum
    (16, 64, 80, 33, 2, None),           # Nnum 16: 256 channels
    (2, 30, 34, 5, 2, None),
    (7, 70, 63, 29, 4, "crop"),
]


def _ids(c):
    return "n%d_X%d_Y%d_M%d_Nz%d_%s" % (c[:5] + (c[5] or "full",))


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_tiled_projector_vs_oracle(env, case, monkeypatch):
    torch, sid_amd = env
    monkeypatch.delenv("SID_TF_T", raising=False)
    n, X, Y, M, Nz, crop = case
    print("case: tiled fft", case, flush=True)
    rng = np.random.default_rng(zlib.crc32(repr(case).encode()) + 5)
    H = random_psf(rng, M, n, Nz)
    CA = _crop(M, Nz) if crop else None
    V = rng.random((X, Y, Nz))
    psf = sid_amd.PSF(H.astype(np.float32), CA, precision="bf16x2", method="fft")
    e_f = rel_l2(psf.forward(V.astype(np.float32)), c_forward(H, V, CA))
    info = psf.info()
    assert info["fft_tile"][0] > 0, info["fft_tile"]
    W = rng.random((X, Y))
    e_b = rel_l2(psf.backward(W.astype(np.float32)), c_backward(H, W, CA))
    print("tile", info["fft_tile"], "rel-L2 forward", e_f, "backward", e_b, flush=True)
    assert e_f < TOL and e_b < TOL
    psf.close()


@pytest.mark.parametrize("T", [16, 24, 32, 36, 40, 48, 64])
def test_tiled_every_length(env, T, monkeypatch):
    """Every instantiated transform length on one geometry (K = 5 taps)."""
    torch, sid_amd = env
    monkeypatch.setenv("SID_TF_T", str(T))
    n, X, Y, M, Nz = 5, 120, 100, 21, 3
    print("case: tiled length", T, flush=True)
    rng = np.random.default_rng(1000 + T)
    H = random_psf(rng, M, n, Nz)
    V = rng.random((X, Y, Nz))
    psf = sid_amd.PSF(H.astype(np.float32), precision="bf16x2", method="fft")
    e_f = rel_l2(psf.forward(V.astype(np.float32)), c_forward(H, V))
    tile = psf.info()["fft_tile"]
    assert tile[0] == T and tile[1] == T - tile[3] + 1, tile
    W = rng.random((X, Y))
    e_b = rel_l2(psf.backward(W.astype(np.float32)), c_backward(H, W))
    print("tile", tile, "rel-L2 forward", e_f, "backward", e_b, flush=True)
    assert e_f < TOL and e_b < TOL
    psf.close()


@pytest.mark.parametrize("X,Y", [(150, 140), (216, 240), (288, 270)], ids=["tiles_20", "tiles_42", "tiles_64"])
def test_tiled_gemm_column_blocks(env, X, Y, monkeypatch):
    """More tiles than one 16-wide GEMM column block: 20, 42 and 64 tiles."""
    torch, sid_amd = env
    monkeypatch.setenv("SID_TF_T", "16")
    n, M, Nz = 3, 13, 2
    print("case: tiled column blocks", X, Y, flush=True)
    rng = np.random.default_rng(X * Y)
    H = random_psf(rng, M, n, Nz)
    V = rng.random((X, Y, Nz))
    psf = sid_amd.PSF(H.astype(np.float32), precision="bf16x2", method="fft")
    e_f = rel_l2(psf.forward(V.astype(np.float32)), c_forward(H, V))
    tile = psf.info()["fft_tile"]
    W = rng.random((X, Y))
    e_b = rel_l2(psf.backward(W.astype(np.float32)), c_backward(H, W))
    print("tile", tile, "rel-L2 forward", e_f, "backward", e_b, flush=True)
    assert tile[2] > 16
    assert e_f < TOL and e_b < TOL
    psf.close()


def test_tiled_fewer_input_planes(env, monkeypatch):
    """A forward input with fewer planes than H (forwardProjectGPU.m:18-19):
    the missing planes contribute nothing, also after a call with all planes."""
    torch, sid_amd = env
    monkeypatch.delenv("SID_TF_T", raising=False)
    n, X, Y, M, Nz = 5, 40, 45, 15, 4
    print("case: tiled nz < Nz", flush=True)
    rng = np.random.default_rng(44)
    H = random_psf(rng, M, n, Nz)
    psf = sid_amd.PSF(H.astype(np.float32), precision="bf16x2", method="fft")
    Vall = rng.random((X, Y, Nz))
    assert rel_l2(psf.forward(Vall.astype(np.float32)), c_forward(H, Vall)) < TOL
    V2 = rng.random((X, Y, 2))
    assert rel_l2(psf.forward(V2.astype(np.float32)), c_forward(H[..., :2], V2)) < TOL
    psf.close()


def test_tiled_mixed_plan_and_isra(env, monkeypatch):
    """Direct and tiled FFT planes in one plan (the halves overlapped on two
    streams), forward, backward and 4 ISRA iterations against the oracle."""
    torch, sid_amd = env
    monkeypatch.delenv("SID_TF_T", raising=False)
    n, X, Y, M, Nz = 5, 55, 50, 21, 4
    print("case: tiled mixed plan", flush=True)
    rng = np.random.default_rng(45)
    H = random_psf(rng, M, n, Nz)
    psf = sid_amd.PSF(H.astype(np.float32), precision="bf16x2")
    psf.set_plane_methods(["direct", "fft", "direct", "fft"])
    V = rng.random((X, Y, Nz))
    O = psf.forward(V.astype(np.float32))
    assert psf.info()["plane_method"] == ["direct", "fft", "direct", "fft"]
    assert psf.info()["fft_tile"][0] > 0
    # the direct planes carry the BF16 PSF (2^-9 relative per weight)
    assert rel_l2(O, c_forward(H, V)) < 2e-3
    W = rng.random((X, Y))
    assert rel_l2(psf.backward(W.astype(np.float32)), c_backward(H, W)) < 2e-3
    htf = c_backward(H, c_forward(H, V))
    got = psf.deconv_rl(htf.astype(np.float32), 4, htf.astype(np.float32))
    want = c_isra(H, htf, htf, 4)
    e = rel_l2(got, want)
    print("isra rel-L2", e, flush=True)
    assert e < 3e-3
    psf.close()


def test_tiled_deterministic(env, monkeypatch):
    """Repeated projections give bit-identical results."""
    torch, sid_amd = env
    monkeypatch.delenv("SID_TF_T", raising=False)
    n, X, Y, M, Nz = 15, 90, 75, 45, 3
    rng = np.random.default_rng(46)
    H = random_psf(rng, M, n, Nz)
    psf = sid_amd.PSF(H.astype(np.float32), precision="bf16x2", method="fft")
    V = rng.random((X, Y, Nz)).astype(np.float32)
    a = psf.forward(V)
    b = psf.forward(V)
    W = rng.random((X, Y)).astype(np.float32)
    c = psf.backward(W)
    d = psf.backward(W)
    assert np.array_equal(a, b) and np.array_equal(c, d)
    psf.close()


def test_tiled_plane_groups(env, monkeypatch):
    """Planes of very different tap ranges split into groups, each with its own
    transform length; forward, backward and ISRA against the oracle, and equal
    (to the FP16 rounding) to the same planes in one group."""
    torch, sid_amd = env
    monkeypatch.delenv("SID_TF_T", raising=False)
    monkeypatch.delenv("SID_TF_GROUPS", raising=False)
    monkeypatch.setenv("SID_TF_GROUP_US", "0")   # groups pay no launch cost in the model
    n, X, Y, M, Nz = 5, 200, 190, 61, 6
    print("case: tiled plane groups", flush=True)
    rng = np.random.default_rng(47)
    H = random_psf(rng, M, n, Nz)
    m = M // 2
    radii = [2, 4, 9, 14, 22, 30]   # lenslet taps 3 .. 13 per axis
    CA = np.array([[m + 1 - r, m + 1 + r] for r in radii], dtype=np.int32)
    V = rng.random((X, Y, Nz))
    W = rng.random((X, Y))
    psf = sid_amd.PSF(H.astype(np.float32), CA, precision="bf16x2", method="fft")
    O = psf.forward(V.astype(np.float32))
    groups = psf.info()["fft_groups"]
    print("groups (planes, T)", groups, flush=True)
    assert len(groups) >= 2 and sum(g[0] for g in groups) == Nz
    assert len(set(g[1] for g in groups)) >= 2
    B = psf.backward(W.astype(np.float32))
    assert rel_l2(O, c_forward(H, V, CA)) < TOL and rel_l2(B, c_backward(H, W, CA)) < TOL
    htf = c_backward(H, c_forward(H, V, CA), CA)
    got = psf.deconv_rl(htf.astype(np.float32), 3, htf.astype(np.float32))
    e = rel_l2(got, c_isra(H, htf, htf, 3, CA))
    print("isra rel-L2", e, flush=True)
    assert e < 4 * TOL
    psf.close()
    monkeypatch.setenv("SID_TF_GROUPS", "1")
    one = sid_amd.PSF(H.astype(np.float32), CA, precision="bf16x2", method="fft")
    O1 = one.forward(V.astype(np.float32))
    assert len(one.info()["fft_groups"]) == 1
    B1 = one.backward(W.astype(np.float32))
    assert rel_l2(O1, O) < TOL and rel_l2(B1, B) < TOL
    one.close()


def test_tiled_two_chains_and_fused_tail(env, monkeypatch):
    """Every plane on the tiled path with several groups: the groups run in two
    stream chains and the backward's inverse transforms apply the ISRA update
    and the column-major outputs themselves.  Repeated calls are bit-identical;
    one chain (sid_psf_set_overlap(0)) and the unfused reduction
    (SID_TF_NOFUSE=1) give the same iterates (bit-identical for the reduction,
    FP32 rounding of the forward image's summation order for the chains); the
    history equals the iterates of separate calls; all match the oracle."""
    torch, sid_amd = env
    monkeypatch.delenv("SID_TF_T", raising=False)
    monkeypatch.delenv("SID_TF_GROUPS", raising=False)
    monkeypatch.delenv("SID_TF_NOFUSE", raising=False)
    monkeypatch.setenv("SID_TF_GROUP_US", "0")
    n, X, Y, M, Nz = 5, 200, 190, 61, 6
    print("case: tiled two chains, fused tail", flush=True)
    rng = np.random.default_rng(48)
    H = random_psf(rng, M, n, Nz)
    m = M // 2
    CA = np.array([[m + 1 - r, m + 1 + r] for r in [2, 4, 9, 14, 22, 30]], dtype=np.int32)
    V = rng.random((X, Y, Nz))
    htf = c_backward(H, c_forward(H, V, CA), CA).astype(np.float32)
    x0 = (htf + 0.05).astype(np.float32)
    psf = sid_amd.PSF(H.astype(np.float32), CA, precision="bf16x2", method="fft")
    a, hist = psf.deconv_rl(htf, 3, x0, history=True)
    assert len(psf.info()["fft_groups"]) >= 2
    b = psf.deconv_rl(htf, 3, x0)
    assert np.array_equal(a, b)
    assert np.array_equal(hist[-1], a)
    assert np.array_equal(hist[0], psf.deconv_rl(htf, 1, x0))
    W = rng.random((X, Y)).astype(np.float32)
    B2 = psf.backward(W)
    want = c_isra(H, htf.astype(np.float64), x0.astype(np.float64), 3, CA)
    e = rel_l2(a, want)
    print("isra rel-L2", e, flush=True)
    assert e < 4 * TOL and rel_l2(B2, c_backward(H, W, CA)) < TOL
    psf.set_overlap(False)
    c = psf.deconv_rl(htf, 3, x0)
    assert rel_l2(c, a) < 1e-6
    psf.close()
    monkeypatch.setenv("SID_TF_NOFUSE", "1")
    unf = sid_amd.PSF(H.astype(np.float32), CA, precision="bf16x2", method="fft")
    d = unf.deconv_rl(htf, 3, x0)
    assert np.array_equal(d, a)
    assert np.array_equal(unf.backward(W), B2)
    unf.close()
