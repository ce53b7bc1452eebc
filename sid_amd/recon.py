"""Single-image reconstruction wiring of external/LFrecon_util/reconstruct_S.m.

Host-side FP64 pre-processing (the reference does it on host arrays as well,
reconstruct_S.m:168-170,191) followed by the device solver (ISRA or
LS_deconv), optionally through a neuron kernel (src/generate_kernel.m).
"""
from __future__ import annotations

import math

import numpy as np


def quantile_type5(x, q: float) -> float:
    """MATLAB ``quantile`` of a vector (Hyndman-Fan type 5), FP64.  Only the
    two order statistics around n*q + 1/2 are selected (no full sort)."""
    x = np.asarray(x, dtype=np.float64).ravel()
    n = x.size
    if n == 0:
        return float("nan")
    pos = n * float(q) + 0.5
    if pos <= 1:
        return float(x.min())
    if pos >= n:
        return float(x.max())
    k = int(math.floor(pos))
    part = np.partition(x, (k - 1, k))
    return float(part[k - 1] + (pos - k) * (part[k] - part[k - 1]))


def _colon(a: float, b: float) -> np.ndarray:
    """MATLAB a:b (unit step)."""
    m = int(math.floor(b - a + 1e-10))
    return a + np.arange(m + 1, dtype=np.float64) if m >= 0 else np.zeros(0)


def fspecial_gaussian(hsize: int, sigma: float) -> np.ndarray:
    """fspecial('gaussian', hsize, sigma) (square hsize): exp(-(x^2+y^2)/(2 sigma^2))
    on a grid centred at (hsize-1)/2, values below eps*max set to 0, sum 1."""
    r = (hsize - 1) / 2.0
    x = np.arange(hsize, dtype=np.float64) - r
    h = np.exp(-(x[:, None] ** 2 + x[None, :] ** 2) / (2.0 * sigma * sigma))
    h[h < np.finfo(np.float64).eps * h.max()] = 0.0
    s = h.sum()
    return h / s if s != 0 else h


def generate_kernel(ker_shape: str, ker_param):
    """src/generate_kernel.m:1-29: 'ball' (ellipsoid indicator of radii
    (p1, p1, p2)), 'gaussian' (fspecial lateral Gaussian of size
    ceil(10 p1) + 1 times exp(-z^2 / (4 p2^2)) for z = -3p2:3p2), 'lorentz'
    (1 / (1 + p1 (x^2 + y^2) + p2 z^2) on ceil(5 p)-radius grids) or 'user'
    (p itself); always divided by its L2 norm.  Returns (kx, ky, kz) float32,
    indexed like MATLAB."""
    if ker_shape == "user":
        k = np.asarray(ker_param, dtype=np.float64)
        if k.ndim == 2:
            k = k[:, :, None]
    else:
        p1, p2 = (float(v) for v in np.asarray(ker_param, dtype=np.float64).ravel()[:2])
        if ker_shape == "ball":
            ax, az = _colon(-p1, p1), _colon(-p2, p2)
            Xg, Yg, Zg = np.meshgrid(ax, ax, az, indexing="ij")
            k = ((Xg ** 2 / p1 ** 2 + Yg ** 2 / p1 ** 2 + Zg ** 2 / p2 ** 2) <= 1).astype(np.float64)
        elif ker_shape == "gaussian":
            hs = int(math.ceil(10 * p1)) + 1
            g = fspecial_gaussian(hs, p1)
            z = _colon(-3 * p2, 3 * p2)
            k = g[:, :, None] * np.exp(-z ** 2 / 4.0 / p2 ** 2)[None, None, :]
        elif ker_shape == "lorentz":
            ax, az = _colon(-math.ceil(5 * p1), math.ceil(5 * p1)), _colon(-math.ceil(5 * p2), math.ceil(5 * p2))
            Xg, Yg, Zg = np.meshgrid(ax, ax, az, indexing="ij")
            k = 1.0 / (1.0 + (p1 * (Xg ** 2 + Yg ** 2) + p2 * Zg ** 2))
        else:
            raise ValueError(f"unknown kernel shape {ker_shape!r}")
    nrm = float(np.linalg.norm(k.ravel()))
    return (k / nrm if nrm > 0 else k).astype(np.float32)


def preprocess(img, microlenses=None, q: float = 0.8):
    """reconstruct_S.m:169-170,191: threshold at the q-quantile of the pixels
    outside the microlens mask, clamp at 0, divide by the max, and return the
    image divided by its L2 norm together with the norm (FP64 throughout)."""
    img = np.asarray(img, dtype=np.float64)
    flat = img.ravel(order="F")
    if microlenses is not None:
        flat = flat[np.asarray(microlenses).ravel(order="F") == 0]
    thr = quantile_type5(flat, q)
    out = np.maximum(img - thr, 0.0)
    mx = out.max()
    out = out / mx
    nrm = math.sqrt(float(np.sum(out * out)))
    return out / nrm, nrm


def reconstruct(psf, img, max_iter: int = 8, microlenses=None, q: float = 0.8, opts=None):
    """reconstruct_S.m:168-205 for one image on the device (the GPU branch):
    Xguess = B(img/nrm); lamb_L1 = min(quantile(Xguess(:), 1 - 1e-4), lamb_L1);
    then whichSolver 'ISRA' (deconvRL(F, B, Xguess, max_iter, Xguess)) or
    'fast_nnls' (LS_deconv(F, B, Xguess, opts)); returns nrm * X (float32,
    MATLAB-indexed (X, Y, Nz)).  ``opts`` keys: whichSolver (default 'ISRA'
    here, the workload's solver), ker_shape / ker_param (neuron kernel,
    reconstruct_S.m:110-128), lamb_L1, lamb_L2, lamb_TV_L2, lamb_TV_L1."""
    opts = dict(opts or {})
    solver = opts.get("whichSolver", "ISRA")
    if "ker_param" in opts:
        psf.set_kernel(generate_kernel(opts.get("ker_shape", "ball"), opts["ker_param"]))
    else:
        psf.set_kernel(None)
    y, nrm = preprocess(img, microlenses, q)
    htf = psf.backward(y.astype(np.float32))
    if solver == "ISRA":
        X = psf.deconv_rl(htf, int(max_iter), htf)
    elif solver == "fast_nnls":
        ls = {k: opts[k] for k in ("lamb_L2", "lamb_TV_L2", "lamb_TV_L1") if k in opts}
        ls["lamb_L1"] = min(quantile_type5(htf, 1 - 1e-4), float(opts.get("lamb_L1", 0.0)))
        ls["max_iter"] = int(max_iter)
        X = psf.ls_deconv(htf, ls)
    else:
        raise ValueError(f"solver {solver!r} is not on the device path")
    return (nrm * X.astype(np.float64)).astype(np.float32)
