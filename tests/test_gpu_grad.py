"""Adjoint kernels: Poisson NLL gradients w.r.t. volume and PSF against the
FP64 oracle, central finite differences, the autograd wrappers, and
bit-exact checkpoint/resume of the refinement loop."""
import numpy as np
import pytest

from oracle import sid_oracle as O
from tests._helpers import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    import sid_amd
    from sid_amd import grad
    if sid_amd.device_count() < 1:
        pytest.fail("no gfx950 device visible")
    return torch, sid_amd, grad


def _case(seed=0, n=5, X=30, Y=25, M=11, nz=3, nb=2):
    rng = np.random.default_rng(seed)
    H = rng.random((M, M, n, n, nz)) + 0.05
    H /= H.sum(axis=(0, 1), keepdims=True)
    V = rng.random((nb, X, Y, nz)) + 0.2
    y = np.stack([rng.poisson(O.forward_project(H, V[f]) * 50.0) / 50.0 for f in range(nb)])
    return H, V, y


def _to_dev(torch, H, V, y):
    Ht = torch.from_numpy(np.ascontiguousarray(H.transpose(4, 3, 2, 1, 0)).astype(np.float32)).cuda()
    Vt = torch.from_numpy(np.ascontiguousarray(V.transpose(0, 3, 2, 1)).astype(np.float32)).cuda()
    yt = torch.from_numpy(np.ascontiguousarray(y.transpose(0, 2, 1)).astype(np.float32)).cuda()
    return Ht, Vt, yt


def _oracle_grads(H, V, y):
    H32 = H.astype(np.float32).astype(np.float64)
    V32 = V.astype(np.float32).astype(np.float64)
    y32 = y.astype(np.float32).astype(np.float64)
    F = lambda v: O.forward_project(H32, v)
    B = lambda o: O.backward_project(H32, o)
    nll = sum(O.poisson_nll(F, V32[f], y32[f]) for f in range(len(V)))
    gV = np.stack([O.poisson_nll_grad_volume(F, B, V32[f], y32[f]) for f in range(len(V))])
    gH = sum(O.poisson_nll_grad_psf(H32, V32[f], y32[f]) for f in range(len(V)))
    return nll, gV, gH


def test_nll_and_gradients_vs_oracle(env):
    torch, sid_amd, grad = env
    print("case: nll grads n5 X30 Y25 M11 Nz3 nb2", flush=True)
    H, V, y = _case()
    Ht, Vt, yt = _to_dev(torch, H, V, y)
    psf = sid_amd.PSF(Ht, full_support=True)
    nll, gV, gH = grad.poisson_nll_grad(psf, Vt, yt)
    nll_o, gV_o, gH_o = _oracle_grads(H, V, y)
    assert abs(nll - nll_o) < 1e-5 * abs(nll_o)
    e_v = rel_l2(gV.cpu().numpy().transpose(0, 3, 2, 1), gV_o)
    e_h = rel_l2(gH.cpu().numpy().transpose(4, 3, 2, 1, 0), gH_o)
    print("rel-L2 gV", e_v, "gH", e_h, flush=True)
    assert e_v < 1e-5 and e_h < 1e-5


@pytest.mark.parametrize("n,X,Y,M,nz,nb", [(3, 30, 27, 7, 2, 3), (15, 45, 30, 31, 2, 2)])
def test_psf_gradient_tiles_vs_oracle(env, n, X, Y, M, nz, nb):
    """The PSF-gradient GEMM at other tile widths: Nnum 3 (one 16-column
    tile, 90 lenslets = three chunks of 32, the last partial, over 3 frames)
    and Nnum 15 (the 15-tile kernel of full-size PSFs)."""
    torch, sid_amd, grad = env
    print(f"case: gH tiles n{n} X{X} Y{Y} M{M} Nz{nz} nb{nb}", flush=True)
    H, V, y = _case(seed=7, n=n, X=X, Y=Y, M=M, nz=nz, nb=nb)
    Ht, Vt, yt = _to_dev(torch, H, V, y)
    psf = sid_amd.PSF(Ht, full_support=True)
    nll, gV, gH = grad.poisson_nll_grad(psf, Vt, yt)
    nll_o, gV_o, gH_o = _oracle_grads(H, V, y)
    assert abs(nll - nll_o) < 1e-5 * abs(nll_o)
    e_v = rel_l2(gV.cpu().numpy().transpose(0, 3, 2, 1), gV_o)
    e_h = rel_l2(gH.cpu().numpy().transpose(4, 3, 2, 1, 0), gH_o)
    print("rel-L2 gV", e_v, "gH", e_h, flush=True)
    assert e_v < 1e-5 and e_h < 1e-5


def test_gradients_vs_central_finite_differences(env):
    """eps = 1e-3 central differences of the FP64 oracle NLL on 64 random
    coordinates (32 of V, 32 of H) against the device gradients (tolerance
    1e-4, relative to max(1, |fd|))."""
    torch, sid_amd, grad = env
    print("case: finite differences n5 X30 Y25 M11 Nz3 nb1", flush=True)
    H, V, y = _case(seed=1, nb=1)
    H = H.astype(np.float32).astype(np.float64)
    V = V.astype(np.float32).astype(np.float64)
    y = y.astype(np.float32).astype(np.float64)
    Ht, Vt, yt = _to_dev(torch, H, V, y)
    psf = sid_amd.PSF(Ht, full_support=True)
    _, gV, gH = grad.poisson_nll_grad(psf, Vt, yt)
    gV = gV.cpu().numpy().transpose(0, 3, 2, 1)[0].astype(np.float64)
    gH = gH.cpu().numpy().transpose(4, 3, 2, 1, 0).astype(np.float64)
    rng = np.random.default_rng(2)
    eps = 1e-3
    nll = lambda HH, VV: O.poisson_nll(lambda v: O.forward_project(HH, v), VV, y[0])
    worst = 0.0
    for _ in range(32):
        i, j, c = rng.integers(0, 30), rng.integers(0, 25), rng.integers(0, 3)
        Vp, Vm = V[0].copy(), V[0].copy()
        Vp[i, j, c] += eps
        Vm[i, j, c] -= eps
        fd = (nll(H, Vp) - nll(H, Vm)) / (2 * eps)
        worst = max(worst, abs(fd - gV[i, j, c]) / max(1.0, abs(fd)))
    for _ in range(32):
        idx = tuple(int(rng.integers(0, s)) for s in H.shape)
        Hp, Hm = H.copy(), H.copy()
        Hp[idx] += eps
        Hm[idx] -= eps
        fd = (nll(Hp, V[0]) - nll(Hm, V[0])) / (2 * eps)
        worst = max(worst, abs(fd - gH[idx]) / max(1.0, abs(fd)))
    print("worst FD relative error", worst, flush=True)
    assert worst < 1e-4


def test_volume_gradient_is_backward_projection(env):
    """dNLL/dV equals the backward projection (backwardProjectGPU_new) of the
    residual image 1 - y/F."""
    torch, sid_amd, grad = env
    print("case: gV == B(1 - y/F)", flush=True)
    H, V, y = _case(seed=3, nb=1)
    Ht, Vt, yt = _to_dev(torch, H, V, y)
    psf = sid_amd.PSF(Ht, full_support=True)
    _, gV, _ = grad.poisson_nll_grad(psf, Vt, yt, True, False)
    F = psf.forward(Vt[0])
    r = 1.0 - yt[0] / F
    B = psf.backward(r)
    assert rel_l2(gV[0].cpu().numpy(), B.cpu().numpy()) < 1e-6


def test_autograd_functions(env):
    torch, sid_amd, grad = env
    print("case: autograd Project / PoissonNLL", flush=True)
    Project, PoissonNLL = grad.functions()
    H, V, y = _case(seed=4)
    Ht, Vt, yt = _to_dev(torch, H, V, y)
    psf = sid_amd.PSF(Ht, full_support=True)
    Vr = Vt.clone().requires_grad_(True)
    Hr = Ht.clone().requires_grad_(True)
    loss = PoissonNLL.apply(Vr, Hr, yt, psf)
    loss.backward()
    nll, gV, gH = grad.poisson_nll_grad(psf, Vt, yt)
    assert abs(loss.item() - nll) < 1e-9 * abs(nll)
    assert torch.equal(Vr.grad, gV) and torch.equal(Hr.grad, gH)
    # Project: d<G, F(V)>/dV = B(G), /dH = project_vjp
    Vr2 = Vt.clone().requires_grad_(True)
    Hr2 = Ht.clone().requires_grad_(True)
    G = torch.rand((Vt.shape[0], Vt.shape[2], Vt.shape[3]), device="cuda")
    out = Project.apply(Vr2, Hr2, psf)
    (out * G).sum().backward()
    gV_o = np.stack([O.backward_project(H, G[f].cpu().numpy().T.astype(np.float64)) for f in range(len(V))])
    assert rel_l2(Vr2.grad.cpu().numpy().transpose(0, 3, 2, 1), gV_o) < 1e-5
    _, gH2 = grad.project_vjp(psf, Vt, G, False, True)
    assert torch.equal(Hr2.grad, gH2)


def test_autograd_follows_a_new_psf_tensor_at_a_reused_address(env):
    """An update loop makes a new H each step (H = H - lr * g); the caching
    allocator often hands it the freed block of an older H.  The handle must
    follow the tensor, not its address."""
    torch, sid_amd, grad = env
    print("case: autograd with H tensors at a reused address", flush=True)
    Project, _ = grad.functions()
    H, V, y = _case(seed=6)
    Ht, Vt, _ = _to_dev(torch, H, V, y)
    psf = sid_amd.PSF(Ht, full_support=True)
    H1 = Ht.clone()
    out1 = Project.apply(Vt, H1, psf)
    ptr = H1.data_ptr()
    del H1
    H2 = torch.empty_like(Ht)
    H2.copy_(Ht * 1.5)
    print("address reused:", H2.data_ptr() == ptr, flush=True)
    out2 = Project.apply(Vt, H2, psf)
    assert rel_l2(out2.cpu().numpy(), 1.5 * out1.cpu().numpy()) < 1e-6
    # an in-place change of the same tensor is seen through its version counter
    H2.mul_(2.0)
    out3 = Project.apply(Vt, H2, psf)
    assert rel_l2(out3.cpu().numpy(), 3.0 * out1.cpu().numpy()) < 1e-6


def test_refinement_checkpoint_resume_bit_exact(env, tmp_path):
    torch, sid_amd, grad = env
    print("case: refinement 20 steps, resume after 10", flush=True)
    H, V, y = _case(seed=5)
    Ht, Vt, yt = _to_dev(torch, H, V, y)
    V0 = torch.full_like(Vt, 0.5)
    psf = sid_amd.PSF(Ht, full_support=True)
    Va, Ha, ha = grad.refine(psf, Ht, V0, yt, 20, checkpoint=str(tmp_path / "a.st"))
    # EM never increases the NLL; 1e-6 relative room for FP32 rounding once it plateaus
    assert len(ha) == 20 and all(ha[i + 1] <= ha[i] + 1e-6 * abs(ha[i]) for i in range(19)), ha
    ck = str(tmp_path / "b.st")
    grad.refine(psf, Ht, V0, yt, 20, checkpoint=ck, stop_after=10)
    Vb, Hb, hb = grad.refine(psf, Ht, V0, yt, 20, checkpoint=ck)
    assert hb == ha
    assert torch.equal(Va, Vb) and torch.equal(Ha, Hb)


@pytest.mark.parametrize("n,X,Y,M,nz,crop", [(5, 33, 27, 11, 3, False), (3, 20, 17, 8, 2, False),
                                             (5, 40, 35, 13, 3, True), (15, 60, 75, 31, 2, False)])
def test_fp64_back_projected_normalisation(env, n, X, Y, M, nz, crop):
    """B(1) = backwardProjectGPU_new(H, ones) in FP64 from summed-area tables, against
    the FP64 oracle's backward projection of a sensor of ones (odd and even M,
    cropped kernels, partial lenslets, an all-zero plane)."""
    torch, sid_amd, grad = env
    print(f"case: FP64 B(1) n{n} X{X} Y{Y} M{M} Nz{nz} crop {crop}", flush=True)
    rng = np.random.default_rng(n * 100 + M)
    H = rng.random((M, M, n, n, nz)) + 0.05
    H[:, :, :, :, nz - 1] = 0.0
    CA = None
    if crop:
        m = M // 2
        CA = np.array([[m + 1 - 2 - c, m + 1 + 2 + c] for c in range(nz)], dtype=np.int32)
    H32 = H.astype(np.float32)
    psf = sid_amd.PSF(H32, CAindex=CA)
    B1 = grad.backward_ones(psf, X, Y).cpu().numpy().transpose(2, 1, 0)
    H64 = H32.astype(np.float64)
    want = O.backward_project(H64, np.ones((X, Y))) if CA is None else O.backward_project_acc(H64, np.ones((X, Y)), CA)
    assert np.all(B1[:, :, nz - 1] == 0.0)
    assert rel_l2(B1, want) < 1e-13
    # interior voxels: the kernel's whole sum (every entry lands on the sensor)
    if CA is None and X > 2 * M and Y > 2 * M:
        q = (X // 2, Y // 2)
        assert abs(B1[q[0], q[1], 0] - H64[:, :, q[0] % n, q[1] % n, 0].sum()) <= 1e-12 * B1[q[0], q[1], 0]
    psf.close()
