"""Host-buffer entry points with buffers large enough (>= 16 MB) to take the
pinned multi-threaded transfer path: results must equal the device entry
points' bit for bit (the transfers only move bytes)."""
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
    return torch, sid_amd


def test_projections_and_isra_host_equal_device(env):
    torch, sid = env
    print("case: host entries n3 X303 Y297 M7 Nz60 (21.6 MB volumes)", flush=True)
    rng = np.random.default_rng(11)
    n, X, Y, M, Nz = 3, 303, 297, 7, 60
    H = random_psf(rng, M, n, Nz).astype(np.float32)
    V = rng.random((X, Y, Nz), dtype=np.float32)
    psf = sid.PSF(H)
    O_host = psf.forward(V)
    Vt = torch.from_numpy(np.ascontiguousarray(V.transpose(2, 1, 0))).cuda()
    O_dev = psf.forward(Vt).cpu().numpy().T
    assert np.array_equal(O_host, O_dev)
    B_host = psf.backward(O_host)
    B_dev = psf.backward(torch.from_numpy(np.ascontiguousarray(O_host.T)).cuda()).cpu().numpy().transpose(2, 1, 0)
    assert B_host.shape == (X, Y, Nz) and np.array_equal(B_host, B_dev)
    X_host = psf.deconv_rl(B_host, 2, B_host)
    Bt = torch.from_numpy(np.ascontiguousarray(B_host.transpose(2, 1, 0))).cuda()
    X_dev = psf.deconv_rl(Bt, 2, Bt).cpu().numpy().transpose(2, 1, 0)
    assert np.array_equal(X_host, X_dev)
    # the source buffer may be reused at once: a second call with other data
    V2 = rng.random((X, Y, Nz), dtype=np.float32)
    assert np.array_equal(psf.forward(V2), psf.forward(torch.from_numpy(np.ascontiguousarray(V2.transpose(2, 1, 0))).cuda()).cpu().numpy().T)
    psf.close()


def test_psf_upload_from_host_large(env):
    torch, sid = env
    print("case: 23 MB PSF uploaded from host vs device", flush=True)
    rng = np.random.default_rng(12)
    n, M, Nz = 15, 31, 27          # 31*31*225*27*4 B = 23.3 MB
    H = random_psf(rng, M, n, Nz).astype(np.float32)
    V = rng.random((90, 75, Nz), dtype=np.float32)
    a = sid.PSF(H)
    b = sid.PSF(torch.from_numpy(np.ascontiguousarray(H.transpose(4, 3, 2, 1, 0))).cuda())
    assert np.array_equal(a.forward(V), b.forward(V))
    a.close()
    b.close()


def test_std_image_host_equal_device(env):
    torch, sid = env
    from sid_amd import movie
    print("case: std image host entry, 120000 px x 48 frames (23 MB)", flush=True)
    rng = np.random.default_rng(13)
    P, T = 120000, 48
    Y = rng.random((P, T), dtype=np.float32)
    s_h, t_h, sd_h, m_h = movie.background_and_std(Y, bg_iter=2)
    s_d, t_d, sd_d, m_d = movie.background_and_std(torch.from_numpy(np.ascontiguousarray(Y.T)).cuda(), bg_iter=2)
    assert m_h == m_d
    assert np.array_equal(sd_h, sd_d.cpu().numpy()) and np.array_equal(s_h, s_d.cpu().numpy())
    assert np.array_equal(t_h, t_d.cpu().numpy())
