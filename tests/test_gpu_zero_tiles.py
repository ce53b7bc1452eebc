"""The BF16 GEMM runs each 32-row weight chunk over the hull of its nonzero
taps only.  Parity with the FP64 oracle where whole chunks are zero, and a PSF
update (which can change which chunks are zero, and the FFT planes' spectra)
gives the same projections as a handle created with the new H."""
import numpy as np
import pytest

from tests._helpers import c_backward, c_forward, random_psf, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sid():
    import sid_amd
    if sid_amd.device_count() < 1:
        pytest.fail("no gfx950 device visible")
    return sid_amd


def _zero_subapertures(H, b_max):
    """Sub-aperture images b < b_max set to zero: forward weight rows ab < n * b_max
    vanish for every tap, so the first 32-row chunk of every tap is all zero."""
    H = H.copy(order="F")
    H[:, :, :, :b_max, :] = 0.0
    return H


@pytest.mark.parametrize("prec,tol", [("bf16x2", 3e-3), ("bf16x3", 2e-5), ("fp32", 2e-6)])
def test_zero_chunks_vs_oracle(sid, prec, tol):
    print("case: zero sub-apertures b < 3, n15 X75 Y90 M37 Nz3,", prec, flush=True)
    rng = np.random.default_rng(71)
    H = _zero_subapertures(random_psf(rng, 37, 15, 3), 3)
    V = rng.random((75, 90, 3))
    psf = sid.PSF(H.astype(np.float32), precision=prec, method="direct")
    O_ref = c_forward(H, V)
    e = rel_l2(psf.forward(V.astype(np.float32)), O_ref)
    print(prec, "forward rel-L2", e, flush=True)
    assert e < tol
    B_ref = c_backward(H, O_ref)
    e = rel_l2(psf.backward(O_ref.astype(np.float32)), B_ref)
    print(prec, "backward rel-L2", e, flush=True)
    assert e < tol
    psf.close()


@pytest.mark.parametrize("prec,method", [("bf16x2", "direct"), ("bf16x2", "fft"), ("fp32", "fft")])
def test_update_equals_fresh_handle(sid, prec, method):
    print("case: update from zero sub-apertures to full H, n15 X90 Y75 M31 Nz3,", prec, method, flush=True)
    rng = np.random.default_rng(72)
    H2 = random_psf(rng, 31, 15, 3).astype(np.float32)
    H1 = _zero_subapertures(H2, 3)
    V = rng.random((90, 75, 3), dtype=np.float32)
    a = sid.PSF(H1, precision=prec, method=method)
    O1 = a.forward(V)
    a.backward(O1)
    a.update(H2)
    b = sid.PSF(H2, precision=prec, method=method)
    O_a, O_b = a.forward(V), b.forward(V)
    assert np.array_equal(O_a, O_b)
    assert not np.array_equal(O_a, O1)
    assert np.array_equal(a.backward(O_b), b.backward(O_b))
    a.close()
    b.close()
