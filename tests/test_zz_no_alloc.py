"""sid_std_image_device_noalloc: the std image without device allocation (its
work space in the library's per-device scratch array) against
sid_std_image_device on the same movies, bit for bit, on two calls in a row
(the second smaller, so the array holds the first call's values), and the
rejection of a movie whose work space exceeds the array."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    import sid_amd
    from sid_amd._lib import lib
    if sid_amd.device_count() < 1:
        pytest.fail("no gfx950 device visible")
    return torch, lib()


def _run(torch, L, name, Y, it, flags):
    P, T = Y.shape
    y = torch.from_numpy(np.ascontiguousarray(Y.T)).cuda()          # column-major [P, T]
    s = torch.full((P,), -1.0, device="cuda")
    t = torch.full((T,), -1.0, device="cuda")
    sd = torch.full((P,), -1.0, device="cuda")
    ymax = ctypes.c_double(0.0)
    torch.cuda.synchronize()
    vp = lambda a: ctypes.c_void_p(a.data_ptr())
    rc = getattr(L, name)(vp(y), P, T, it, flags, vp(s), vp(t), vp(sd), ctypes.byref(ymax), None)
    assert rc == 0, L.sid_last_error()
    return s.cpu().numpy(), t.cpu().numpy(), sd.cpu().numpy(), ymax.value


@pytest.mark.parametrize("it,flags", [(2, 1), (0, 0)])
def test_noalloc_matches_the_allocating_entry(env, it, flags):
    torch, L = env
    rng = np.random.default_rng(7 + it)
    for P, T in ((300_000, 120), (20_011, 37)):                     # second call smaller
        print("case: std image no-alloc P=%d T=%d bg_iter=%d flags=%d" % (P, T, it, flags), flush=True)
        Y = (rng.random((P, T)) * 2 + np.outer(rng.random(P), 1 + np.cos(np.arange(T) / 7.0))).astype(np.float32)
        a = _run(torch, L, "sid_std_image_device", Y, it, flags)
        b = _run(torch, L, "sid_std_image_device_noalloc", Y, it, flags)
        for u, v in zip(a[:3], b[:3]):
            assert np.array_equal(u, v)
        assert a[3] == b[3]


def test_noalloc_rejects_a_movie_beyond_its_scratch(env):
    torch, L = env
    y = torch.zeros(16, device="cuda")
    ymax = ctypes.c_double(0.0)
    vp = lambda a: ctypes.c_void_p(a.data_ptr())
    # 4 M pixels need about 96 MiB of work space (24 B per pixel)
    rc = L.sid_std_image_device_noalloc(vp(y), 4_000_000, 2, 1, 0, None, None, vp(y), ctypes.byref(ymax), None)
    assert rc == 2 and b"work space" in L.sid_last_error()
