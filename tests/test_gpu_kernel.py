"""Neuron-kernel wrapping (reconstruct_S.m:122-128,184-190): the device
convn(V, k, 'same') against the oracle, and the wrapped projectors, ISRA and
LS_deconv against the oracle composed the same way
(F(convn(X, k)), convn(B(y), k))."""
import numpy as np
import pytest

from oracle import sid_oracle as O
from tests._helpers import c_backward, c_forward, random_psf, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sid():
    import sid_amd
    if sid_amd.device_count() < 1:
        pytest.fail("no gfx950 device visible")
    return sid_amd


CONV_CASES = [
    ((40, 35, 6), (3, 3, 3)),
    ((70, 19, 4), (17, 17, 5)),      # the find_kernel-sized neuron, wider than one 64-px tile edge
    ((33, 41, 3), (4, 2, 6)),        # even sides, kernel deeper than the volume
    ((20, 20, 1), (5, 7, 1)),        # one plane
    ((131, 17, 2), (65, 1, 1)),      # the largest lateral side
    ((90, 50, 4), (33, 9, 3)),       # a two-step Toeplitz window
    ((70, 80, 3), (3, 65, 2)),       # the tallest kernel: B fragments staged in several chunks
]


@pytest.mark.parametrize("form", ["fft", "direct"])
@pytest.mark.parametrize("vd,kd", CONV_CASES, ids=lambda c: "x".join(map(str, c)))
def test_convn_same_vs_oracle(sid, vd, kd, form, monkeypatch):
    """FP32 convn: the transform form (default where it fits) and the direct
    FP32 sum (SID_CONVN_FFT=0), each within 1e-6 rel-L2 of the FP64 oracle."""
    import torch
    if form == "direct":
        monkeypatch.setenv("SID_CONVN_FFT", "0")
    print("case: convn", form, vd, kd, flush=True)
    rng = np.random.default_rng(sum(vd) * 7 + sum(kd))
    V = rng.random(vd).astype(np.float32)
    k = rng.random(kd).astype(np.float32)
    ref = O.convn_same(V.astype(np.float64), k.astype(np.float64))
    got = sid.convn_same(V, k)
    e = rel_l2(got, ref)
    print("convn rel-L2", e, flush=True)
    assert e < 1e-6
    # device tensors in the reversed shape give the same values
    Vt = torch.from_numpy(np.ascontiguousarray(V.transpose(2, 1, 0))).cuda()
    kt = torch.from_numpy(np.ascontiguousarray(k.transpose(2, 1, 0))).cuda()
    gt = sid.convn_same(Vt, kt).cpu().numpy().transpose(2, 1, 0)
    assert np.array_equal(gt, got)


@pytest.mark.parametrize("prec,bar", [("bf16x2", 2.0 ** -9), ("bf16x3", 2.0 ** -16)])
@pytest.mark.parametrize("vd,kd", CONV_CASES, ids=lambda c: "x".join(map(str, c)))
def test_convn_mfma_vs_oracle(sid, vd, kd, prec, bar, monkeypatch):
    """BF16-split MFMA convn (Toeplitz form, forced by SID_CONVN_FFT=0): the
    volume is carried as hi + lo BF16 (2^-17 relative), the kernel as BF16
    (2^-9 per tap) in BF16X2 or hi + lo in BF16X3; the bars are those
    per-element bounds."""
    monkeypatch.setenv("SID_CONVN_FFT", "0")
    print("case: convn mfma", prec, vd, kd, flush=True)
    rng = np.random.default_rng(sum(vd) * 11 + sum(kd))
    V = rng.random(vd).astype(np.float32)
    k = rng.random(kd).astype(np.float32)
    ref = O.convn_same(V.astype(np.float64), k.astype(np.float64))
    got = sid.convn_same(V, k, precision=prec)
    e = rel_l2(got, ref)
    print("convn", prec, "rel-L2", e, flush=True)
    assert e < bar


FFT_CASES = CONV_CASES + [
    ((3, 2, 2), (65, 1, 1)),         # kernel far wider than the volume
    ((300, 260, 7), (17, 17, 5)),    # 512-point transforms, the full-size neuron kernel
    ((1000, 40, 3), (9, 5, 3)),      # 1024-point x transform (32 lanes x 32 points)
    ((30, 1000, 3), (3, 9, 3)),      # 1024-point y transform
    ((700, 30, 2), (5, 5, 1)),       # 768-point x transform (16 lanes x 48 points)
    ((1100, 20, 2), (9, 5, 3)),      # 2048-point x transform (32 lanes x 64 points)
    ((20, 1100, 2), (3, 9, 2)),      # 2048-point y transform
    ((50, 40, 10), (5, 5, 8)),       # the deepest kernel of the transform form
    ((45, 37, 9), (6, 4, 7)),        # even kernel sides, 128-point transforms
    ((1, 1, 1), (1, 1, 1)),          # one voxel, one tap
    ((7, 1, 3), (3, 1, 3)),          # one row: 15 of the 16 rows of an x block are padding
    ((5, 3, 2), (2, 2, 2)),          # even 2-tap kernel on a tiny volume
]


@pytest.mark.parametrize("prec", ["fp32", "bf16x2", "bf16x3"])
@pytest.mark.parametrize("vd,kd", FFT_CASES, ids=lambda c: "x".join(map(str, c)))
def test_convn_fft_vs_oracle(sid, vd, kd, prec):
    """Transform form of convn (the default at every precision where it fits,
    transform lengths up to 2048): FP32 2-D DFTs in (x, y) with the z sum
    between them.  FP32 rounding in the
    transforms (log2 N stages) stays below 2e-6 rel-L2 of the FP64 direct sum."""
    print("case: convn fft", prec, vd, kd, flush=True)
    rng = np.random.default_rng(sum(vd) * 13 + sum(kd))
    V = rng.random(vd).astype(np.float32)
    k = rng.random(kd).astype(np.float32)
    ref = O.convn_same(V.astype(np.float64), k.astype(np.float64))
    got = sid.convn_same(V, k, precision=prec)
    e = rel_l2(got, ref)
    print("convn fft", prec, "rel-L2", e, flush=True)
    assert e < 2e-6


def test_convn_fft_nonfinite(sid):
    """Transform form with one NaN voxel: every (x, y) of the output planes the
    kernel's depth reaches is NaN (an FFT convolution spreads it along the
    transformed axes); the other planes are finite and match the oracle."""
    print("case: convn fft nonfinite", flush=True)
    rng = np.random.default_rng(21)
    V = rng.random((40, 35, 9)).astype(np.float32)
    k = rng.random((3, 5, 3)).astype(np.float32)
    V[5, 7, 4] = np.nan
    got = sid.convn_same(V, k, precision="bf16x2")
    assert np.isnan(got[:, :, 3:6]).all()
    ref = O.convn_same(np.nan_to_num(V, nan=0.0).astype(np.float64), k.astype(np.float64))
    far = [0, 1, 2, 6, 7, 8]
    assert np.isfinite(got[:, :, far]).all()
    assert rel_l2(got[:, :, far], ref[:, :, far]) < 2e-6


def test_convn_fft_reuse(sid):
    """The handle-free entry's work arrays are reused across calls: another
    kernel on the same volume, then a smaller volume, each against the oracle."""
    print("case: convn fft reuse", flush=True)
    rng = np.random.default_rng(3)
    for vd, kd in [((80, 70, 6), (9, 9, 3)), ((80, 70, 6), (5, 3, 5)), ((31, 17, 2), (9, 9, 3))]:
        V = rng.random(vd).astype(np.float32)
        k = rng.random(kd).astype(np.float32)
        ref = O.convn_same(V.astype(np.float64), k.astype(np.float64))
        assert rel_l2(sid.convn_same(V, k, precision="bf16x2"), ref) < 2e-6


@pytest.mark.parametrize("form", ["fft", "mfma"])
def test_wrapped_projectors_bf16x2(sid, form, monkeypatch):
    """A BF16X2 handle with a kernel (convn in the transform or the MFMA form):
    projections within the mode's budget; a second kernel replaces the first
    (the handle's kernel spectra are rebuilt)."""
    if form == "mfma":
        monkeypatch.setenv("SID_CONVN_FFT", "0")
    print("case: wrapped projectors bf16x2", form, flush=True)
    rng, H, k, F, B = _wrapped_case(seed=6)
    psf = sid.PSF(H.astype(np.float32), precision="bf16x2")
    psf.set_kernel(k.astype(np.float32))
    V = rng.random((40, 35, 4))
    W = rng.random((40, 35))
    e_f = rel_l2(psf.forward(V.astype(np.float32)), F(V))
    e_b = rel_l2(psf.backward(W.astype(np.float32)), B(W))
    print("wrapped bf16x2 forward", e_f, "backward", e_b, flush=True)
    assert e_f < 2e-3 and e_b < 2e-3
    k2 = rng.random((3, 7, 2))
    k2 = (k2 / np.linalg.norm(k2)).astype(np.float32).astype(np.float64)
    psf.set_kernel(k2.astype(np.float32))
    e2 = rel_l2(psf.forward(V.astype(np.float32)), c_forward(H, O.convn_same(V, k2)))
    print("second kernel forward", e2, flush=True)
    assert e2 < 2e-3


def _wrapped_case(seed=5, n=5, X=40, Y=35, M=15, nz=4, kd=(5, 5, 3)):
    rng = np.random.default_rng(seed)
    H = random_psf(rng, M, n, nz).astype(np.float32).astype(np.float64)
    k = rng.random(kd)
    k = (k / np.linalg.norm(k)).astype(np.float32).astype(np.float64)
    F = lambda v: c_forward(H, O.convn_same(v, k))
    B = lambda o: O.convn_same(c_backward(H, o), k)
    return rng, H, k, F, B


@pytest.mark.parametrize("form", ["fft", "direct"])
def test_wrapped_projectors_vs_oracle(sid, form, monkeypatch):
    if form == "direct":
        monkeypatch.setenv("SID_CONVN_FFT", "0")
    print("case: wrapped projectors n5 X40 Y35 M15 Nz4 k5x5x3", form, flush=True)
    rng, H, k, F, B = _wrapped_case()
    psf = sid.PSF(H.astype(np.float32))
    psf.set_kernel(k.astype(np.float32))
    V = rng.random((40, 35, 4))
    W = rng.random((40, 35))
    e_f = rel_l2(psf.forward(V.astype(np.float32)), F(V))
    e_b = rel_l2(psf.backward(W.astype(np.float32)), B(W))
    print("wrapped forward", e_f, "backward", e_b, flush=True)
    assert e_f < 5e-6 and e_b < 5e-6
    # removing the kernel restores the plain projectors
    psf.set_kernel(None)
    assert rel_l2(psf.forward(V.astype(np.float32)), c_forward(H, V)) < 5e-6


def test_wrapped_projectors_new_volume_size(sid):
    """One handle with a kernel, two volume sizes in turn (the transform form
    re-plans its lengths and keeps its larger arrays), then the first again."""
    print("case: wrapped projectors, volume size changes", flush=True)
    rng, H, k, F, B = _wrapped_case(seed=12)
    psf = sid.PSF(H.astype(np.float32))
    psf.set_kernel(k.astype(np.float32))
    for shape in [(40, 35, 4), (30, 26, 4), (40, 35, 4)]:
        V = rng.random(shape)
        got = psf.forward(V.astype(np.float32))
        assert rel_l2(got, c_forward(H, O.convn_same(V, k))) < 5e-6, shape


def test_wrapped_isra_and_ls_vs_oracle(sid):
    print("case: wrapped ISRA / LS n5 X40 Y35 M15 Nz4 k5x5x3", flush=True)
    rng, H, k, F, B = _wrapped_case(seed=8)
    psf = sid.PSF(H.astype(np.float32))
    psf.set_kernel(k.astype(np.float32))
    y = F(rng.random((40, 35, 4)))
    htf = B(y).astype(np.float32).astype(np.float64)
    X_ref = O.deconv_rl(F, B, htf, 4, htf)
    X, hist = psf.deconv_rl(htf.astype(np.float32), 4, htf.astype(np.float32), history=True)
    e_r = rel_l2(X, X_ref)
    assert np.array_equal(hist[-1], X)
    X_ls_ref = O.ls_deconv(F, B, htf, max_iter=3, lamb_L1=0.01)
    X_ls = psf.ls_deconv(htf.astype(np.float32), {"max_iter": 3, "lamb_L1": 0.01})
    e_l = rel_l2(X_ls, X_ls_ref)
    print("wrapped isra", e_r, "ls", e_l, flush=True)
    assert e_r < 2e-5 and e_l < 1e-4


@pytest.mark.parametrize("form", ["fft", "mfma"])
def test_wrapped_isra_bf16x2_vs_oracle(sid, form, monkeypatch):
    """deconvRL on a BF16X2 handle with a kernel (convn in the transform or the
    MFMA form), against the oracle's FP64 loop composed the same way; the bar is
    the BF16X2 projector's budget over 4 iterations."""
    if form == "mfma":
        monkeypatch.setenv("SID_CONVN_FFT", "0")
    print("case: wrapped ISRA bf16x2", form, flush=True)
    rng, H, k, F, B = _wrapped_case(seed=9)
    psf = sid.PSF(H.astype(np.float32), precision="bf16x2")
    psf.set_kernel(k.astype(np.float32))
    y = F(rng.random((40, 35, 4)))
    htf = B(y).astype(np.float32).astype(np.float64)
    X_ref = O.deconv_rl(F, B, htf, 4, htf)
    X = psf.deconv_rl(htf.astype(np.float32), 4, htf.astype(np.float32))
    e = rel_l2(X, X_ref)
    print("wrapped isra bf16x2", form, e, flush=True)
    assert e < 3e-3


def test_reconstruct_with_kernel_and_solvers(sid):
    """reconstruct_S GPU branch: Htf = B(img/nrm) (wrapped), lamb_L1 clamp to
    the 1 - 1e-4 quantile of Htf, ISRA or fast_nnls, times nrm."""
    from sid_amd import recon
    print("case: reconstruct ball kernel n5 X40 Y35 M15 Nz4", flush=True)
    rng = np.random.default_rng(12)
    H = random_psf(rng, 15, 5, 4).astype(np.float32).astype(np.float64)
    kern = recon.generate_kernel("ball", (2, 1)).astype(np.float64)
    F = lambda v: c_forward(H, O.convn_same(v, kern))
    B = lambda o: O.convn_same(c_backward(H, o), kern)
    img = F(rng.random((40, 35, 4)))
    y, nrm = recon.preprocess(img)
    htf = B(y.astype(np.float32).astype(np.float64))
    psf = sid.PSF(H.astype(np.float32))
    opts = {"ker_shape": "ball", "ker_param": (2, 1)}
    out = recon.reconstruct(psf, img, max_iter=3, opts=dict(opts, whichSolver="ISRA"))
    ref = nrm * O.deconv_rl(F, B, htf, 3, htf)
    e1 = rel_l2(out, ref)
    lam = min(recon.quantile_type5(htf, 1 - 1e-4), 0.1)
    out2 = recon.reconstruct(psf, img, max_iter=3, opts=dict(opts, whichSolver="fast_nnls", lamb_L1=0.1))
    ref2 = nrm * O.ls_deconv(F, B, htf, max_iter=3, lamb_L1=lam)
    e2 = rel_l2(out2, ref2)
    print("reconstruct isra", e1, "fast_nnls", e2, flush=True)
    assert e1 < 2e-5 and e2 < 1e-4


def test_kernel_errors(sid):
    rng = np.random.default_rng(1)
    H = random_psf(rng, 7, 3, 2).astype(np.float32)
    psf = sid.PSF(H)
    with pytest.raises(sid.SidError, match="SID_ERR_DIMS"):
        psf.set_kernel(np.ones((66, 3, 1), np.float32))
    psf.set_kernel(np.ones((3, 3, 1), np.float32))
    import torch
    from sid_amd import grad
    V = torch.rand((1, 2, 12, 12), device="cuda")
    y = torch.rand((1, 12, 12), device="cuda")
    with pytest.raises(sid.SidError, match="SID_ERR_ARG"):
        grad.poisson_nll_grad(psf, V, y)
