"""Single-image reconstruction wiring of external/LFrecon_util/reconstruct_S.m.

Host-side FP64 pre-processing (the reference does it on host arrays as well,
reconstruct_S.m:168-170,191) followed by the device ISRA loop.
"""
from __future__ import annotations

import math

import numpy as np


def quantile_type5(x, q: float) -> float:
    """MATLAB ``quantile`` of a vector (Hyndman-Fan type 5), FP64."""
    x = np.sort(np.asarray(x, dtype=np.float64).ravel())
    n = x.size
    if n == 0:
        return float("nan")
    pos = n * float(q) + 0.5
    if pos <= 1:
        return float(x[0])
    if pos >= n:
        return float(x[-1])
    k = int(math.floor(pos))
    return float(x[k - 1] + (pos - k) * (x[k] - x[k - 1]))


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


def reconstruct(psf, img, max_iter: int = 8, microlenses=None, q: float = 0.8):
    """reconstruct_S.m:168-205 for one image with whichSolver 'ISRA':
    Xguess = B(img/nrm); X = deconvRL(F, B, Xguess, max_iter, Xguess);
    returns nrm * X (float32, MATLAB-indexed (X, Y, Nz))."""
    y, nrm = preprocess(img, microlenses, q)
    htf = psf.backward(y.astype(np.float32))
    X = psf.deconv_rl(htf, int(max_iter), htf)
    return (nrm * X.astype(np.float64)).astype(np.float32)
