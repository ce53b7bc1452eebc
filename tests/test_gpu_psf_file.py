"""A handle built from a PSF MAT-file projects exactly like one built from the
same H in memory (GPU path without crops, and with the file's CAindex)."""
import os

import numpy as np
import pytest

from tests._helpers import c_forward, rel_l2

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "h5")


@pytest.mark.parametrize("name", ["psf_v73_contiguous.mat", "psf_v73_chunked.mat", "psf_latest.mat"])
def test_psf_from_file(name):
    import sid_amd
    if sid_amd.device_count() < 1:
        pytest.fail("no gfx950 device visible")
    print("case: PSF.from_file", name, flush=True)
    path = os.path.join(GOLD, name)
    exp = np.load(path + ".npz")
    H = np.asfortranarray(exp["H"].T.astype(np.float32))
    ca = np.asarray(exp["CAindex"].T, dtype=np.int32)
    M, n, nz = H.shape[0], H.shape[2], H.shape[4]
    rng = np.random.default_rng(4)
    V = rng.random((4 * n, 5 * n, nz)).astype(np.float32)
    for use_ca in (False, True):
        a = sid_amd.PSF.from_file(path, use_caindex=use_ca)
        b = sid_amd.PSF(H, ca if use_ca else None)
        fa, fb = a.forward(V), b.forward(V)
        assert np.array_equal(fa, fb)
        assert rel_l2(fa, c_forward(H.astype(np.float64), V.astype(np.float64), ca if use_ca else None)) < 2e-6
        a.close()
        b.close()
