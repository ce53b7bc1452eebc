"""Movie background and standard-deviation image on the device: the producer
of the reconstruction input in sid_main.m:88-103 (src/rank_1_factorization.m,
src/compute_std_image.m), on libsidproj.so.

``Y`` is the movie as a matrix of pixels x frames, indexed like MATLAB
(numpy array (P, T)), or a device tensor in the reversed shape (T, P) (the
same column-major memory).
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, lib


def background_and_std(Y, bg_iter: int = 2, bg_sub: bool = True, normalize: bool = True):
    """sid_main.m:88-103: optionally divide the movie by its maximum, the rank-1
    background [bg_spatial, bg_temporal] = rank_1_factorization(Y, bg_iter)
    (zeros when bg_sub is off), and std_image = compute_std_image(Y, bg_spatial,
    bg_temporal).  Returns (bg_spatial (P,), bg_temporal (T,), std_image (P,),
    movie maximum (0.0 when not normalising)); numpy in, numpy out, or device
    tensors in and out."""
    it = int(bg_iter) if bg_sub else 0
    flags = 1 if normalize else 0
    ymax = ctypes.c_double(0.0)
    if type(Y).__module__.startswith("torch"):
        import torch
        y = Y.detach().to(torch.float32).contiguous()
        T, P = y.shape
        s = torch.empty(P, dtype=torch.float32, device=y.device)
        t = torch.empty(T, dtype=torch.float32, device=y.device)
        sd = torch.empty(P, dtype=torch.float32, device=y.device)
        check(lib().sid_std_image_device(ctypes.c_void_p(y.data_ptr()), P, T, it, flags, ctypes.c_void_p(s.data_ptr()),
                                         ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(sd.data_ptr()),
                                         ctypes.byref(ymax),
                                         ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return s, t, sd, ymax.value
    y = np.asfortranarray(np.asarray(Y, dtype=np.float32))
    P, T = y.shape
    s = np.zeros(P, np.float32)
    t = np.zeros(T, np.float32)
    sd = np.zeros(P, np.float32)
    P_ = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    check(lib().sid_std_image(P_(y), P, T, it, flags, P_(s), P_(t), P_(sd), ctypes.byref(ymax)))
    return s, t, sd, ymax.value


def rank_1_factorization(Y, maxIter: int = 1):
    """src/rank_1_factorization.m:1 — [bg_spatial, bg_temporal] of the movie
    (no normalisation), as (P,) and (T,) arrays."""
    s, t, _, _ = background_and_std(Y, bg_iter=maxIter, bg_sub=True, normalize=False)
    return s, t
