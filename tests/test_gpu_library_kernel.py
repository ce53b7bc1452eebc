"""The kernel option of generate_LFM_library_GPU (src/generate_LFM_library_GPU.m:
54-56,95-101) on the device against the FP64 oracle: each seed's volume (the
box average of the listed reconstructions, zero elsewhere in the image_size x
z0..z1 volume) convolved with the kernel by convn 'same' and projected with
planes z0..z1 of H.  Covers boxes clipped at the sensor edges, sub-volumes at
both ends of the PSF's planes, a ball kernel from generate_kernel and an
even-sized user kernel (MATLAB's 'same' window starts at floor(size(k)/2))."""
import numpy as np
import pytest

import oracle.sid_oracle as O
from tests._helpers import random_psf, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import sid_amd
    from sid_amd import library, recon
    if sid_amd.device_count() < 1:
        pytest.fail("no gfx950 device visible")
    return sid_amd, library, recon


CENTRES = np.array([[2.4, 20.5, 1.2],      # x box clipped at 1, first planes
                    [44.5, 38.6, 6.0],     # x and y boxes clipped at the far edges, last planes
                    [22.5, 17.4, 3.5],     # interior
                    [30.2, 2.6, 4.8]])     # y clipped at 1
NEUR_ID = np.array([[1, 0, 1], [0, 1, 0], [1, 1, 1], [0, 0, 1]])


@pytest.mark.parametrize("kind", ["ball", "user"])
def test_library_kernel_vs_oracle(env, kind):
    sid_amd, library, recon_mod = env
    print(f"case: library with kernel {kind}, n5 M21 X45 Y40 Nz6, 4 seeds", flush=True)
    rng = np.random.default_rng(11)
    n, M, X, Y, nz = 5, 21, 45, 40, 6
    H = random_psf(rng, M, n, nz).astype(np.float32)
    recon = [rng.random((X, Y, nz)).astype(np.float32) for _ in range(3)]
    opts = {"image_size": np.array([X, Y]), "axial": 2.0, "neur_rad": 3}
    if kind == "ball":
        opts["ker_shape"] = "ball"
        opts["ker_param"] = [2.0, 1.0]
    else:
        opts["ker_shape"] = "user"
        opts["ker_param"] = rng.random((4, 3, 2))
    kern = recon_mod.generate_kernel(opts["ker_shape"], opts["ker_param"])
    psf = sid_amd.PSF(H)
    fm = library.generate_LFM_library_GPU(recon, CENTRES, NEUR_ID, psf, opts)
    psf.close()
    assert fm.shape == (len(CENTRES), X * Y) and fm.dtype == np.float32
    H64 = H.astype(np.float64)
    for s in range(len(CENTRES)):
        ids = list(np.flatnonzero(NEUR_ID[s]))
        want, _ = O.library_gpu_entry([r.astype(np.float64) for r in recon], ids, CENTRES[s], 3, 2.0, H64,
                                      kernel=kern.astype(np.float64))
        err = rel_l2(fm[s], want.ravel(order="F"))
        print(f"seed {s}: rel-L2 {err:.2e}", flush=True)
        assert err < 2e-6


def test_convn_seeds_vs_oracle(env):
    """The grown boxes and their values: convn 'same' of each box (zero outside)
    on the whole sensor, read on the grown box."""
    sid_amd, library, recon_mod = env
    print("case: seed convolution, 40 boxes on 60 x 50, kernel 5 x 4 x 3", flush=True)
    rng = np.random.default_rng(12)
    X, Y, nz = 60, 50, 7
    k = rng.random((5, 4, 3))
    boxes, vals = [], []
    for _ in range(40):
        c = (rng.uniform(1, X), rng.uniform(1, Y), rng.uniform(1, nz))
        b = library.placement(c, int(rng.integers(1, 6)), int(rng.integers(0, 3)), (X, Y), nz)
        boxes.append(b)
        vals.append(rng.random((b[1] - b[0] + 1, b[3] - b[2] + 1, b[5] - b[4] + 1)).astype(np.float32))
    grown, out = library.convn_seeds(boxes, vals, (X, Y), k.astype(np.float32))
    for s, b in enumerate(boxes):
        V = np.zeros((X, Y, b[5] - b[4] + 1))
        V[b[0] - 1:b[1], b[2] - 1:b[3], :] = vals[s]
        C = O.convn_same(V, k)
        g = grown[s]
        assert (g[4], g[5]) == (b[4], b[5])
        # the grown box holds every nonzero of the convolution
        mask = np.zeros((X, Y), dtype=bool)
        mask[g[0] - 1:g[1], g[2] - 1:g[3]] = True
        assert np.abs(C[~mask]).max(initial=0.0) == 0.0
        got = out[s]
        want = C[g[0] - 1:g[1], g[2] - 1:g[3], :]
        assert got.shape == want.shape
        assert rel_l2(got, want) < 2e-6
