"""deconvRL's guard cases and non-finite inputs on the device, for every plane
method (all planes on the direct GEMM, all on the FFT path, and a mix).

deconvRL.m:85-87: R = Htf ./ D; R(isinf(R)) = 0; X = X .* R; X(isnan(X)) = 0.
D = 0 happens where a voxel's kernel is all zero: x/0 gives Inf (-> 0) and 0/0
gives NaN (X * NaN -> 0), so such voxels end at exactly 0.

Non-finite inputs follow forwardProjectGPU / backwardProjectGPU_new: their FFT
convolutions (conv2FFT_.m:14) turn one NaN or Inf anywhere in the input into
NaN over the whole output, whichever plane it sits on; deconvRL then zeroes the
whole iterate.  The numpy oracle restates those FFT convolutions, so it shows
the same spread.
"""
import zlib

import numpy as np
import pytest

import oracle.sid_oracle as O
from tests._helpers import c_backward, c_forward, c_isra, random_psf, rel_l2

pytestmark = pytest.mark.gpu

N, X, Y, M, NZ = 5, 40, 35, 11, 3
PLANS = ["direct", "fft", "mixed"]


@pytest.fixture(scope="module")
def env():
    import sid_amd
    if sid_amd.device_count() < 1:
        pytest.fail("no gfx950 device visible")
    return sid_amd


def _psf(sid_amd, H, plan, precision="fp32"):
    psf = sid_amd.PSF(H.astype(np.float32), precision=precision)
    if plan in ("direct", "fft"):
        psf.set_method(plan)
    else:
        empty = [np.all(H[:, :, :, :, c] == 0) for c in range(H.shape[4])]
        psf.set_plane_methods(["auto" if empty[c] else ("fft" if c % 2 == 0 else "direct")
                               for c in range(H.shape[4])])
    return psf


def _case(tag):
    rng = np.random.default_rng(zlib.crc32(tag.encode()))
    H = random_psf(rng, M, N, NZ)
    X0 = rng.random((X, Y, NZ)) + 0.1
    return rng, H, X0


@pytest.mark.parametrize("plan", PLANS)
@pytest.mark.parametrize("precision", ["fp32", "bf16x2"])
def test_zero_kernel_guards(env, plan, precision):
    """One all-zero kernel (a, b, c) = (1, 2, 1): its voxels get D = 0.  Where
    Htf > 0 the ratio is Inf -> 0; where Htf = 0 it is NaN and X * NaN -> 0."""
    print("case: zero kernel", plan, precision, flush=True)
    rng, H, X0 = _case("zero-kernel")
    H[:, :, 1, 2, 1] = 0.0
    htf = c_backward(H, c_forward(H, rng.random((X, Y, NZ))))
    sub = np.zeros((X, Y, NZ), dtype=bool)
    sub[1::N, 2::N, 1] = True
    hz = htf.copy()
    hz[sub & (rng.random((X, Y, NZ)) < 0.5)] = 0.0          # 0/0 on half of them
    hz[sub & (hz > 0)] += 1.0                                 # x/0 on the rest
    want = c_isra(H, hz, X0, 3)
    assert np.all(want[sub] == 0.0)
    psf = _psf(env, H, plan, precision)
    got = psf.deconv_rl(hz.astype(np.float32), 3, X0.astype(np.float32))
    psf.close()
    assert np.all(got[sub] == 0.0)
    assert np.all(np.isfinite(got))
    assert rel_l2(got, want) < TOL[precision]


# rel-L2 bars of finite results: FP32 against the FP64 oracle, and the BF16X2
# mode (FFT planes: FP16 spectra, 2^-11 per stored value; direct planes: the
# BF16 PSF, 2^-9) with a margin for the few-iteration ISRA
TOL = {"fp32": 1e-5, "bf16x2": 5e-3}


@pytest.mark.parametrize("plan", PLANS)
@pytest.mark.parametrize("precision", ["fp32", "bf16x2"])
def test_all_zero_plane(env, plan, precision):
    """An all-zero PSF plane: D = 0 on the whole plane, which ends at 0."""
    print("case: zero plane", plan, precision, flush=True)
    rng, H, X0 = _case("zero-plane")
    H[:, :, :, :, 1] = 0.0
    htf = rng.random((X, Y, NZ))
    want = c_isra(H, htf, X0, 2)
    assert np.all(want[:, :, 1] == 0.0)
    psf = _psf(env, H, plan, precision)
    assert psf.info()["plane_method"][1] == "empty"
    got = psf.deconv_rl(htf.astype(np.float32), 2, X0.astype(np.float32))
    psf.close()
    assert np.all(got[:, :, 1] == 0.0)
    assert rel_l2(got, want) < TOL[precision]


@pytest.mark.parametrize("plan", PLANS)
@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
@pytest.mark.parametrize("precision", ["fp32", "bf16x2"])
def test_nonfinite_projector_inputs(env, plan, bad, precision):
    """One non-finite voxel (on a direct or an FFT plane) makes the whole image
    NaN; one non-finite pixel makes the whole back-projection NaN.  BF16X2 with
    every plane on the FFT path writes the back-projection from its inverse
    transforms (no separate reduction)."""
    print("case: non-finite input", plan, bad, precision, flush=True)
    rng, H, X0 = _case("nonfinite")
    psf = _psf(env, H, plan, precision)
    tol = 2e-6 if precision == "fp32" else 2.0 ** -11 if plan == "fft" else 2e-3
    for c in range(NZ):
        V = X0.copy()
        V[7, 9, c] = bad
        assert np.isnan(O.forward_project(H, V)).all()
        img = psf.forward(V.astype(np.float32))
        assert np.isnan(img).all(), (c, np.isnan(img).mean())
    img = psf.forward(X0.astype(np.float32))               # a finite input afterwards is unaffected
    assert rel_l2(img, c_forward(H, X0)) < tol
    o = rng.random((X, Y))
    o[3, 30] = bad
    assert np.isnan(O.backward_project(H, o)).all()
    B = psf.backward(o.astype(np.float32))
    assert np.isnan(B).all()
    o[3, 30] = 0.5
    B = psf.backward(o.astype(np.float32))
    assert rel_l2(B, c_backward(H, o)) < tol
    psf.close()


@pytest.mark.parametrize("plan", PLANS)
@pytest.mark.parametrize("precision", ["fp32", "bf16x2"])
def test_nan_voxel_in_xguess_zeroes_the_reconstruction(env, plan, precision):
    """A NaN voxel in Xguess: F(X) is NaN everywhere, so D and R are too and
    X(isnan(X)) = 0 leaves every voxel at 0 (numpy oracle of the FFT path)."""
    print("case: NaN in Xguess", plan, precision, flush=True)
    rng, H, X0 = _case("nan-xguess")
    htf = c_backward(H, c_forward(H, X0))
    Xg = X0.copy()
    Xg[20, 11, 2] = np.nan
    F = lambda v: O.forward_project(H, v)
    B = lambda o: O.backward_project(H, o)
    with np.errstate(invalid="ignore"):
        want = O.deconv_rl(F, B, htf, 2, Xg)
    assert np.all(want == 0.0)
    psf = _psf(env, H, plan, precision)
    got = psf.deconv_rl(htf.astype(np.float32), 2, Xg.astype(np.float32))
    assert np.all(got == 0.0)
    # the next call on the same handle starts clean
    got = psf.deconv_rl(htf.astype(np.float32), 2, X0.astype(np.float32))
    assert rel_l2(got, c_isra(H, htf, X0, 2)) < TOL[precision]
    psf.close()
