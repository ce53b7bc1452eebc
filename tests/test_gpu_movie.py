"""Movie background (rank 1) and std image on the device against the FP64
oracle (src/rank_1_factorization.m, src/compute_std_image.m, sid_main.m:88-103)."""
import numpy as np
import pytest

from oracle import sid_oracle as O
from tests._helpers import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def movie():
    import sid_amd
    from sid_amd import movie as mv
    if sid_amd.device_count() < 1:
        pytest.fail("no gfx950 device visible")
    return sid_amd, mv


CASES = [(2000, 300, 2, True), (2000, 301, 1, False), (4099, 17, 5, True), (1000, 2, 2, True), (777, 64, 0, True)]


@pytest.mark.parametrize("P,T,it,norm", CASES, ids=lambda c: str(c))
def test_background_and_std_vs_oracle(movie, P, T, it, norm):
    sid_amd, mv = movie
    print("case: std image", P, T, it, norm, flush=True)
    rng = np.random.default_rng(P + T + it)
    Y = (rng.random((P, T)) * 3 + np.outer(rng.random(P), 1 + np.sin(np.arange(T) / 5.0))).astype(np.float32)
    s, t, sd, ymax = mv.background_and_std(Y, bg_iter=it, bg_sub=it > 0, normalize=norm)
    Y64 = Y.astype(np.float64)
    if norm:
        assert ymax == float(Y.max())
        Y64 = Y64 / float(Y.max())
    else:
        assert ymax == 0.0
    if it > 0:
        s_ref, t_ref = O.rank_1_factorization(Y64, it)
    else:
        s_ref, t_ref = np.zeros(P), np.zeros(T)
    std_ref = O.compute_std_image(Y64, s_ref, t_ref)
    e = [rel_l2(sd, std_ref)]
    if it > 0:
        e += [rel_l2(s, s_ref), rel_l2(t, t_ref)]
    else:
        assert not s.any() and not t.any()
    print("std/s/t rel-L2", e, flush=True)
    assert max(e) < 1e-6


def test_device_tensors_match_host_arrays(movie):
    import torch
    sid_amd, mv = movie
    rng = np.random.default_rng(3)
    Y = rng.random((3000, 50)).astype(np.float32)
    host = mv.background_and_std(Y, bg_iter=2)
    dev = mv.background_and_std(torch.from_numpy(np.ascontiguousarray(Y.T)).cuda(), bg_iter=2)
    for a, b in zip(host[:3], dev[:3]):
        assert np.array_equal(a, b.cpu().numpy())
    assert host[3] == dev[3]
    s, t = mv.rank_1_factorization(Y, 2)
    assert np.array_equal(s, mv.background_and_std(Y, 2, normalize=False)[0])


def test_movie_errors(movie):
    sid_amd, mv = movie
    with pytest.raises(sid_amd.SidError, match="SID_ERR_ARG"):
        mv.background_and_std(np.zeros((10, 4), np.float32), bg_iter=1, normalize=True)
    with pytest.raises(sid_amd.SidError, match="SID_ERR_ARG"):
        mv.background_and_std(np.ones((10, 4), np.float32), bg_iter=-1)


@pytest.mark.parametrize("where", ["first", "last", "middle"])
def test_normalising_maximum_at_the_edges(movie, where):
    """The maximum comes out of the first column pass when background
    iterations run, out of a pass of its own otherwise: both see every value."""
    sid_amd, mv = movie
    print("case: movie maximum at", where, flush=True)
    rng = np.random.default_rng(9)
    P, T = 2053, 39                        # ragged against the 4-frame groups and the pixel chunks
    Y = (rng.random((P, T)) * 0.5).astype(np.float32)
    i, j = {"first": (0, 0), "last": (P - 1, T - 1), "middle": (P // 2, T // 2)}[where]
    Y[i, j] = 7.0
    for it in (0, 1, 2):
        assert mv.background_and_std(Y, bg_iter=it, bg_sub=it > 0, normalize=True)[3] == 7.0
