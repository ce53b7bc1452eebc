"""Stream order of the device-pointer entry points.

A NULL stream argument is the device's default stream: the call must see
every write already queued there.  torch's default stream passes NULL, and the
handle's own stream is non-blocking, so without an explicit wait a projection
could read an input that a queued torch kernel has not written yet.  Each case
delays the producer of the input behind a long spin on the default stream and
compares with the result computed after a full synchronisation.
"""
import numpy as np
import pytest

from tests._helpers import random_psf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    import sid_amd
    if sid_amd.device_count() < 1:
        pytest.fail("no gfx950 device visible")
    if not hasattr(torch.cuda, "_sleep"):
        pytest.skip("torch.cuda._sleep is not available")
    return torch, sid_amd


def _delayed(torch, src):
    """A copy of src written on the default stream behind a long spin."""
    dst = torch.zeros_like(src)
    torch.cuda.synchronize()
    torch.cuda._sleep(200_000_000)          # ~0.1 s of spinning on the default stream
    dst.copy_(src)
    return dst


def test_null_stream_inputs_are_ordered(env):
    torch, sid_amd = env
    from sid_amd import grad
    print("case: default-stream ordering of forward / backward / deconvRL / NLL", flush=True)
    assert torch.cuda.current_stream().cuda_stream == 0
    rng = np.random.default_rng(31)
    H = random_psf(rng, 11, 5, 3)
    psf = sid_amd.PSF(H.astype(np.float32), full_support=True)
    V = torch.from_numpy(rng.random((3, 35, 40)).astype(np.float32)).cuda()
    O = torch.from_numpy(rng.random((35, 40)).astype(np.float32)).cuda()
    want_f = psf.forward(V)
    want_b = psf.backward(O)
    h = psf.backward(want_f)
    want_rl = psf.deconv_rl(h, 2, h)
    y = (want_f * 1.3).unsqueeze(0).contiguous()
    want_nll, want_gv, _ = grad.poisson_nll_grad(psf, V.unsqueeze(0), y, True, False)
    torch.cuda.synchronize()
    assert torch.equal(psf.forward(_delayed(torch, V)), want_f)
    assert torch.equal(psf.backward(_delayed(torch, O)), want_b)
    assert torch.equal(psf.deconv_rl(_delayed(torch, h), 2, _delayed(torch, h)), want_rl)
    nll, gv, _ = grad.poisson_nll_grad(psf, _delayed(torch, V.unsqueeze(0)), _delayed(torch, y), True, False)
    assert nll == want_nll and torch.equal(gv, want_gv)
    psf.close()
