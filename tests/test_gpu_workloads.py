"""The standard workloads at their full sizes against the oracle.

* rl_small (255 x 255 px, Nnum 15, 11 planes, 105 px kernels, 40 balls,
  Poisson image at a 1000-photon peak): reconstruct_S pre-processing and 8
  ISRA iterations in the device's FP64 and FP32 arithmetic against the C
  oracle's FP64 ISRA, iterate by iterate.
* library (2500 seeds, 1005 x 1005 px, 51 planes, 345 px kernels): every
  seed's box and patch rectangle integer-identical to the oracle's placement
  (generate_LFM_library_GPU.m:68-88) and project_forward_patched rectangle;
  24 seeds' patches against the oracle's projection of their sub-volume.
* grad (16 frames, 495 x 495 px, 31 planes, 225 px kernels): the FP32 NLL
  gradients against central finite differences (eps 1e-3) and the exact FP64
  derivative on 64 coordinates,
  dNLL/dV against the backward projection of 1 - y/F, and a 20-step
  refinement resumed bit-exactly from its step-10 checkpoint.
"""

import numpy as np
import pytest

import oracle.sid_oracle as O
from tests._helpers import c_backward, c_forward, c_forward_h32, c_isra, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    import sid_amd
    if sid_amd.device_count() < 1:
        pytest.fail("no gfx950 device visible")
    return torch, sid_amd


# --------------------------------------------------------------------------- rl_small

def test_rl_small_isra_vs_oracle(env):
    torch, sid_amd = env
    from sid_amd import recon, synth
    wl = synth.WORKLOADS["rl_small"]
    print("case: rl_small 255x255x11 M105, 8 ISRA iterations, fp64 / fp32 / bf16x2 vs the C oracle", flush=True)
    H, CA = synth.make_psf(wl.n, wl.M, wl.zs)                      # float32, MATLAB layout
    H64 = H.astype(np.float64)
    V = synth.make_volume(wl.X, wl.Y, wl.nz, wl.n_balls, wl.r_lo, wl.r_hi, wl.seed, wl.background)
    img = synth.poisson_image(c_forward(H64, V, CA), wl.peak_photons, wl.seed)
    y, nrm = recon.preprocess(img)
    y_o, nrm_o = O.reconstruct_preprocess(img)
    assert np.abs(y - y_o).max() <= 1e-15 and abs(nrm - nrm_o) <= 1e-12 * nrm_o
    y32 = y.astype(np.float32)
    # oracle: Htf = B(y); X_0 = Htf; 8 iterates (one ISRA step at a time is the same loop)
    htf = c_backward(H64, y32.astype(np.float64), CA)
    Xo, X = [], htf.copy()
    for _ in range(wl.iters):
        X = c_isra(H64, htf, X, 1, CA)
        Xo.append(X)
    psf = sid_amd.PSF(H)                                          # forwardProjectGPU semantics (no crop)
    errs = {}
    for prec in ("fp64", "fp32", "bf16x2"):
        psf.set_precision(prec)
        h = psf.backward(y32)
        X, hist = psf.deconv_rl(h, wl.iters, h, history=True)
        errs[prec] = [rel_l2(hist[i], Xo[i]) for i in range(wl.iters)]
        assert np.array_equal(hist[-1], X)
        print(prec, "rel-L2 per iteration:", ["%.2e" % e for e in errs[prec]], flush=True)
    assert max(errs["fp64"]) < 1e-6            # float32 I/O of the FP64 path
    assert max(errs["fp32"]) < 1e-5
    assert errs["bf16x2"][-1] < 1e-3
    # the reconstruct_S wiring as a whole: nrm * X
    psf.set_precision("fp32")
    rec = recon.reconstruct(psf, img, wl.iters)
    assert rel_l2(rec, nrm_o * Xo[-1]) < 1e-5
    psf.close()


# --------------------------------------------------------------------------- library

def test_library_workload_vs_oracle(env):
    torch, sid_amd = env
    from sid_amd import library, synth
    wl = synth.WORKLOADS["library"]
    print("case: library 2500 seeds at 1005x1005x51, M345", flush=True)
    dev = torch.device("cuda:0")
    Ht, CA = synth.make_psf(wl.n, wl.M, wl.zs, xp=torch, device=dev)   # (Nz, n, n, M, M)
    psf = sid_amd.PSF(Ht)
    sd = synth.make_seeds(wl)
    boxes, vals = library.workload_seeds(wl)
    ns = len(boxes)
    # placement: generate_LFM_library_GPU.m:68-88 with dz = ceil(r / axial); axial = r / (dz - 1/2) gives dz
    for k in range(ns):
        r, dz = int(sd["box"][k]) // 2, int(sd["depth"][k]) // 2
        p = O.library_gpu_placement(sd["center"][k], r, r / (dz - 0.5), (wl.X, wl.Y), wl.nz)
        assert tuple(boxes[k]) == (p["x0"], p["x1"], p["y0"], p["y1"], p["z0"], p["z1"]), k
    lib = library.Library(psf, (wl.X, wl.Y))
    rects, offs = lib.rects(boxes)
    for k in range(ns):
        b = boxes[k]
        g = O.patched_rect((b[0], b[2], b[4]), (b[1], b[3], b[5]), (wl.X, wl.Y), wl.n, wl.M)
        assert tuple(rects[k]) == (g["x0"], g["x1"], g["y0"], g["y1"]), k
        assert offs[k + 1] - offs[k] == (g["x1"] - g["x0"] + 1) * (g["y1"] - g["y0"] + 1)
    b, v, voff, out, rects, offs = lib.prepare(boxes, vals)
    lib.project(b, v, voff, out, use_graph=True)
    torch.cuda.synchronize()
    # 24 seeds spread over the batch: the oracle projects each sub-volume on a
    # sensor window aligned to the lenslet grid that holds its box and rectangle
    n = wl.n
    worst = 0.0
    for s in np.linspace(0, ns - 1, 24).astype(int):
        x0, x1, y0, y1, z0, z1 = (int(t) for t in boxes[s])
        r = rects[s]
        wx0 = (min(x0, r[0]) - 1) // n * n + 1
        wy0 = (min(y0, r[2]) - 1) // n * n + 1
        wx1, wy1 = max(x1, r[1]), max(y1, r[3])
        Vw = np.zeros((wx1 - wx0 + 1, wy1 - wy0 + 1, z1 - z0 + 1))
        Vw[x0 - wx0:x1 - wx0 + 1, y0 - wy0:y1 - wy0 + 1, :] = vals[s]
        Hs = Ht[z0 - 1:z1].cpu().numpy().transpose(4, 3, 2, 1, 0)
        full = c_forward_h32(Hs, Vw, CA[z0 - 1:z1], threads=16)
        ref = full[r[0] - wx0:r[1] - wx0 + 1, r[2] - wy0:r[3] - wy0 + 1]
        got = library.Library.patch(out, rects, offs, s)
        e = rel_l2(got, ref)
        worst = max(worst, e)
        # the rectangle holds the seed's whole footprint within the sensor
        assert abs(full.sum() - ref.sum()) <= 1e-9 * full.sum()
    print("library: worst of 24 seeds rel-L2 %.2e" % worst, flush=True)
    assert worst < 2e-6
    del Ht
    psf.close()


# --------------------------------------------------------------------------- grad

def _grad_workload():
    """16 seeded ball volumes and a PSF that carries the photon scale (unit-sum
    kernels times the gain that makes the brightest clean image peak at the
    workload's photon count).  The gradients are taken at 60-80 % of the true
    volumes, where the residual 1 - y/F has one sign over the lit pixels, so the
    sums are not dominated by cancellation."""
    from sid_amd import synth
    wl = synth.WORKLOADS["grad"]
    H, _ = synth.make_psf(wl.n, wl.M, wl.zs)
    Vs, V0 = [], []
    rng = np.random.default_rng(wl.seed)
    for f in range(wl.batch):
        V = synth.make_volume(wl.X, wl.Y, wl.nz, wl.n_balls, wl.r_lo, wl.r_hi, wl.seed + 17 * f, wl.background)
        Vs.append(V)
        V0.append((V * rng.uniform(0.6, 0.8, V.shape)).astype(np.float32))
    return wl, H, Vs, V0


def test_grad_workload_fd_adjoint_and_resume(env, tmp_path):
    torch, sid_amd = env
    from sid_amd import grad
    print("case: grad 16 frames 495x495x31 M225: FD on 64 coordinates, gV = B(1 - y/F), 20-step resume", flush=True)
    wl, H, Vs, V0 = _grad_workload()
    n, X, Y, nz, nb, M = wl.n, wl.X, wl.Y, wl.nz, wl.batch, wl.M
    off = M // 2
    dev = torch.device("cuda:0")
    tv = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.float32).transpose(2, 1, 0))).to(dev)
    psf = sid_amd.PSF(H, full_support=True)
    # clean images from the direct FP64 path: exactly 0 where the kernels' truncated
    # zeros leave a pixel dark (FFT-path planes would leave rounding noise there)
    psf.set_precision("fp64")
    clean = [psf.forward(tv(v)).cpu().numpy().T.astype(np.float64) for v in Vs]
    gain = wl.peak_photons / max(float(c.max()) for c in clean)
    H = (H * np.float32(gain)).astype(np.float32, order="F")
    psf.update(H)
    ys = [np.random.default_rng(1000 + f).poisson(clean[f] * gain).astype(np.float32) for f in range(nb)]
    Ht = torch.from_numpy(np.ascontiguousarray(H.transpose(4, 3, 2, 1, 0))).to(dev)
    Vt = torch.stack([tv(v) for v in V0])                         # (nb, Nz, Y, X)
    yt = torch.stack([torch.from_numpy(np.ascontiguousarray(y.T)).to(dev) for y in ys])
    # F = forward(V0) per frame in the FP64 device arithmetic, then the FP32 gradients
    F = [psf.forward(Vt[f]).cpu().numpy().T.astype(np.float64) for f in range(nb)]
    psf.set_precision("fp32")
    nll, gV, gH = grad.poisson_nll_grad(psf, Vt, yt)
    gV = gV.cpu().numpy()                                         # (nb, Nz, Y, X)
    gH = gH.cpu().numpy()                                         # (Nz, n, n, M, M)
    y64 = [y.astype(np.float64) for y in ys]
    V64 = [v.astype(np.float64) for v in V0]
    R = [np.where(y64[f] > 0, 1.0 - y64[f] / np.where(y64[f] > 0, F[f], 1.0), 1.0) for f in range(nb)]
    eps = 1e-3

    def nll_diff(f, sl, dF):
        # NLL(+eps) - NLL(-eps) of frame f where F changes by +-dF on window sl (F is
        # linear in V and in H); y log F is 0 where y = 0 (there F may be 0)
        Fw, yw = F[f][sl], y64[f][sl]
        m = yw > 0
        assert np.all(Fw[m] - np.abs(dF[m]) > 0)
        lg = np.zeros_like(Fw)
        lg[m] = np.log1p(dF[m] / Fw[m]) - np.log1p(-dF[m] / Fw[m])
        return float(np.sum(2.0 * dF - yw * lg))

    rng = np.random.default_rng(99)
    worst_fd = worst_an = 0.0
    for _ in range(32):                                           # volume coordinates: dF = eps * kernel at the voxel
        f, i, j, c = int(rng.integers(nb)), int(rng.integers(X)), int(rng.integers(Y)), int(rng.integers(nz))
        K = H[:, :, i % n, j % n, c].astype(np.float64)
        pi0, pi1 = max(0, i - off), min(X, i - off + M)
        pj0, pj1 = max(0, j - off), min(Y, j - off + M)
        Kw = K[pi0 - i + off:pi1 - i + off, pj0 - j + off:pj1 - j + off]
        sl = (slice(pi0, pi1), slice(pj0, pj1))
        fd = nll_diff(f, sl, eps * Kw) / (2 * eps)
        an = float(np.sum(Kw * R[f][sl]))                         # the exact derivative
        g = float(gV[f, c, j, i])
        e_fd, e_an = abs(fd - g) / max(1.0, abs(fd)), abs(an - g) / max(1.0, abs(an))
        worst_fd, worst_an = max(worst_fd, e_fd), max(worst_an, e_an)
        if e_fd > 1e-4:
            print("  V(f=%d, %d, %d, %d): fd %.8e analytic %.8e grad %.8e" % (f, i, j, c, fd, an, g), flush=True)
    for _ in range(32):                                           # PSF coordinates: dF = eps * shifted sublattice plane
        a, b, c = int(rng.integers(n)), int(rng.integers(n)), int(rng.integers(nz))
        i, j = off + int(rng.integers(-40, 41)), off + int(rng.integers(-40, 41))
        dx, dy = i - off, j - off
        xs, xd = (slice(0, X - dx), slice(dx, X)) if dx >= 0 else (slice(-dx, X), slice(0, X + dx))
        ys_, yd = (slice(0, Y - dy), slice(dy, Y)) if dy >= 0 else (slice(-dy, Y), slice(0, Y + dy))
        tot = an = 0.0
        for f in range(nb):
            S = np.zeros((X, Y))
            S[a::n, b::n] = V64[f][a::n, b::n, c]
            Sd = np.zeros((X, Y))
            Sd[xd, yd] = S[xs, ys_]
            tot += nll_diff(f, (slice(None), slice(None)), eps * Sd)
            an += float(np.sum(Sd * R[f]))
        fd = tot / (2 * eps)
        g = float(gH[c, b, a, j, i])
        e_fd, e_an = abs(fd - g) / max(1.0, abs(fd)), abs(an - g) / max(1.0, abs(an))
        worst_fd, worst_an = max(worst_fd, e_fd), max(worst_an, e_an)
        if e_fd > 1e-4:
            print("  H(%d, %d, %d, %d, %d): fd %.8e analytic %.8e grad %.8e" % (i, j, a, b, c, fd, an, g), flush=True)
    print("grad: worst relative error over 64 coordinates: FD %.2e, analytic %.2e" % (worst_fd, worst_an), flush=True)
    assert worst_fd < 1e-4
    assert worst_an < 1e-5
    # dNLL/dV = B(1 - y/F) (backwardProjectGPU_new of the residual; 1 where y = 0),
    # both on the direct plan the gradient runs
    psf.set_method("direct")
    for f in (0, nb - 1):
        Ff = psf.forward(Vt[f])
        r = torch.where(yt[f] > 0, 1.0 - yt[f] / torch.where(yt[f] > 0, Ff, torch.ones_like(Ff)), torch.ones_like(Ff))
        Bf = psf.backward(r).cpu().numpy()
        assert rel_l2(gV[f], Bf) < 1e-6
    # 20 refinement steps; the same run stopped after 10 and resumed is bit-identical
    ck_b = str(tmp_path / "b.st")
    Va, Ha, ha = grad.refine(psf, Ht, Vt, yt, 20)
    assert len(ha) == 20 and all(ha[k + 1] <= ha[k] + 1e-6 * abs(ha[k]) for k in range(19)), ha
    # (a checkpoint every 10 steps: each one holds the 1.9 GB of V and H)
    grad.refine(psf, Ht, Vt, yt, 20, checkpoint=ck_b, stop_after=10, checkpoint_every=10)
    Vb, Hb, hb = grad.refine(psf, Ht, Vt, yt, 20, checkpoint=ck_b, checkpoint_every=10)
    assert hb == ha
    assert torch.equal(Va, Vb) and torch.equal(Ha, Hb)
    print("grad: NLL %.6e -> %.6e over 20 steps" % (ha[0], ha[-1]), flush=True)
    psf.close()
