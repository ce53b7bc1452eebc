"""Batched library projection against the FP64 oracle: each seed's patch is
the full-sensor forward projection (generate_LFM_library_GPU semantics) of its
sub-volume with planes z0..z1 of H, on project_forward_patched's rectangle."""
import numpy as np
import pytest

from tests._helpers import c_forward, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    import sid_amd
    from sid_amd import library
    if sid_amd.device_count() < 1:
        pytest.fail("no gfx950 device visible")
    return torch, sid_amd, library


def _seeds(rng, ns, X, Y, nz):
    boxes, vals = [], []
    for _ in range(ns):
        c = (rng.uniform(1, X), rng.uniform(1, Y), rng.uniform(1, nz))
        r = int(rng.integers(1, 6))
        dz = int(rng.integers(0, 3))
        from sid_amd.library import placement, gaussian_blob
        b = placement(c, r, dz, (X, Y), nz)
        boxes.append(b)
        vals.append(gaussian_blob(b, c, 2.0) + np.float32(0.01))
    return np.asarray(boxes, dtype=np.int64), vals


@pytest.mark.parametrize("n,M,X,Y,nz,ns", [(5, 21, 70, 65, 6, 30), (15, 61, 120, 105, 4, 12)])
def test_library_vs_oracle(env, n, M, X, Y, nz, ns):
    torch, sid_amd, library = env
    print(f"case: library n{n} M{M} X{X} Y{Y} Nz{nz} seeds{ns}", flush=True)
    rng = np.random.default_rng(n * 1000 + M)
    H = rng.random((M, M, n, n, nz)) * (rng.random((M, M, 1, 1, nz)) > 0.3)
    psf = sid_amd.PSF(H.astype(np.float32))
    lib = library.Library(psf, (X, Y))
    boxes, vals = _seeds(rng, ns, X, Y, nz)
    b, v, voff, out, rects, offs = lib.prepare(boxes, vals)
    lib.project(b, v, voff, out, use_graph=True)
    first = out.clone()
    lib.project(b, v, voff, out, use_graph=True)       # graph replay
    second = out.clone()
    lib.project(b, v, voff, out, use_graph=False)      # plain launches
    third = out.clone()
    assert rel_l2(second.cpu().numpy(), first.cpu().numpy()) < 1e-6
    assert rel_l2(third.cpu().numpy(), first.cpu().numpy()) < 1e-6
    worst = 0.0
    for s in range(ns):
        x0, x1, y0, y1, z0, z1 = (int(t) for t in boxes[s])
        V = np.zeros((X, Y, z1 - z0 + 1))
        V[x0 - 1:x1, y0 - 1:y1, :] = vals[s]
        full = c_forward(H[:, :, :, :, z0 - 1:z1].astype(np.float32).astype(np.float64), V)
        r = rects[s]
        ref = full[r[0] - 1:r[1], r[2] - 1:r[3]]
        got = library.Library.patch(first, rects, offs, s)
        e = rel_l2(got, ref)
        worst = max(worst, e)
        # the rectangle holds the whole footprint: nothing is lost outside it
        assert abs(np.abs(full).sum() - np.abs(ref).sum()) <= 1e-9 * np.abs(full).sum()
    print("worst seed rel-L2", worst, flush=True)
    assert worst < 2e-6
