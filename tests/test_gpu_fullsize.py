"""The precision target at full size: 8 ISRA iterations on the 1005 x 1005 x 51
workload (M up to 345), BF16X2 (autotuned hybrid of the split MFMA GEMM and
the FP16-spectra FFT path) within 1e-3 relative L2 of the device FP64 path
after the 8th iterate, FP32 within 1e-5.  The FP64 path itself is pinned to
the oracle by tests/test_gpu_projector.py."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_rl_full_precision_targets():
    import torch
    import sid_amd
    from sid_amd import recon, synth
    if sid_amd.device_count() < 1:
        pytest.fail("no gfx950 device visible")
    print("case: rl_full 1005x1005x51, 8 ISRA iterations, fp64 / fp32 / bf16x2", flush=True)
    wl = synth.WORKLOADS["rl_full"]
    dev = torch.device("cuda:0")
    H, _ = synth.make_psf(wl.n, wl.M, wl.zs, xp=torch, device=dev)
    psf = sid_amd.PSF(H)
    del H
    V = synth.make_volume(wl.X, wl.Y, wl.nz, wl.n_balls, wl.r_lo, wl.r_hi, wl.seed, wl.background)
    Vt = torch.from_numpy(np.ascontiguousarray(V.transpose(2, 1, 0))).to(dev)
    img = synth.poisson_image(psf.forward(Vt).cpu().numpy().T, wl.peak_photons, wl.seed)
    y, _ = recon.preprocess(img)
    yt = torch.from_numpy(np.ascontiguousarray(y.T.astype(np.float32))).to(dev)
    iters = 8
    psf.set_precision("fp64")
    h64 = psf.backward(yt)
    X64 = psf.deconv_rl(h64, iters, h64).double()
    errs = {}
    for prec in ("fp32", "bf16x2"):
        psf.set_precision(prec)
        h = psf.backward(yt)
        X = psf.deconv_rl(h, iters, h).double()
        errs[prec] = float((X - X64).norm() / X64.norm())
        del h, X
    print("rel-L2 after 8 iterations vs FP64:", errs, flush=True)
    assert errs["fp32"] < 1e-5
    assert errs["bf16x2"] < 1e-3
    psf.close()
