"""Normalised footprint library (src/generate_LFM_library_CPU.m) on the device
against the FP64 oracle: ball voxelisation (MATLAB colon / ind2sub / round),
the kernel path, clipping at the volume's borders, project_forward_patched's
rectangle, Q / norm(Q) and the sparse rows of find(Q)."""
import numpy as np
import pytest

import oracle.sid_oracle as O
from tests._helpers import random_psf, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import sid_amd
    from sid_amd import library
    if sid_amd.device_count() < 1:
        pytest.fail("no gfx950 device visible")
    return sid_amd, library


def _compare(library, psf, H, centers, r, dim, m, n, kernel=None):
    row_ptr, cols, vals = library.generate_LFM_library_CPU(centers, psf, r, dim, m, kernel)
    assert row_ptr.shape == (len(centers) + 1,)
    worst = 0.0
    for s, c in enumerate(centers):
        got_c = cols[row_ptr[s]:row_ptr[s + 1]]
        got_v = vals[row_ptr[s]:row_ptr[s + 1]]
        try:
            Q, _ = O.library_cpu_entry(c, r, dim, m, H, n, kernel)
        except ValueError:            # no voxel inside the volume: an empty row here
            assert got_c.size == 0
            continue
        assert np.all(np.diff(got_c) > 0)
        dense = np.zeros(m[0] * m[1])
        dense[got_c] = got_v
        assert np.array_equal(got_c, np.flatnonzero(Q))
        e = rel_l2(dense, Q)
        worst = max(worst, e)
        assert abs(np.linalg.norm(got_v) - 1.0) < 1e-12
    print("worst row rel-L2 %.2e" % worst, flush=True)
    assert worst < 2e-6


def test_ball_library_vs_oracle(env):
    sid_amd, library = env
    n, M, m = 5, 21, (60, 55, 6)
    print("case: footprint library, ball r=2 rr=4/3, n5 M21", flush=True)
    rng = np.random.default_rng(41)
    H = random_psf(rng, M, n, m[2])
    psf = sid_amd.PSF(H.astype(np.float32))
    centers = np.array([[30.2, 27.5, 3.0], [1.4, 10.0, 2.0], [59.6, 54.5, 5.5], [12.0, 40.0, 0.6],
                        [-20.0, 5.0, 2.0], [33.5, 2.5, 6.4]])
    _compare(library, psf, H.astype(np.float32).astype(np.float64), centers, 2, (1.0, 1.0, 1.5), m, n)
    psf.close()


def test_kernel_library_vs_oracle(env):
    sid_amd, library = env
    n, M, m = 3, 9, (30, 33, 4)
    print("case: footprint library, 3x3x3 kernel, n3 M9", flush=True)
    rng = np.random.default_rng(42)
    H = random_psf(rng, M, n, m[2])
    ker = (rng.random((3, 3, 3)) + 0.2).astype(np.float32)
    ker[0, 2, :] = 0
    psf = sid_amd.PSF(H.astype(np.float32))
    centers = np.array([[15.0, 16.0, 2.0], [2.4, 31.6, 3.5], [28.0, 1.0, 1.0]])
    _compare(library, psf, H.astype(np.float32).astype(np.float64), centers, None, None, m, n,
             kernel=ker.astype(np.float64))
    psf.close()


def test_ball_library_nnum15_synthetic_psf(env):
    """Nnum 15 with the synthetic PSF, whose truncated kernels leave exact zeros
    inside the rectangle (both sides must agree on them)."""
    sid_amd, library = env
    from sid_amd import synth
    n, M, zs = 15, 61, (-2, -1, 0, 1, 2)
    print("case: footprint library, ball r=3 rr=1.5, n15 M61 synthetic PSF", flush=True)
    H, _ = synth.make_psf(n, M, zs)
    m = (120, 105, len(zs))
    psf = sid_amd.PSF(H)
    centers = np.array([[60.3, 52.5, 3.0], [7.0, 99.6, 1.2], [118.5, 4.4, 4.6]])
    _compare(library, psf, H.astype(np.float64), centers, 3, (1.0, 1.0, 2.0), m, n)
    psf.close()


def test_library_nsf_rejects_even_geometry(env):
    sid_amd, library = env
    rng = np.random.default_rng(43)
    H = random_psf(rng, 10, 3, 2)
    psf = sid_amd.PSF(H.astype(np.float32))
    with pytest.raises(RuntimeError):
        library.generate_LFM_library_CPU(np.array([[5.0, 5.0, 1.0]]), psf, 1, (1.0, 1.0, 1.0), (20, 20))
    psf.close()
