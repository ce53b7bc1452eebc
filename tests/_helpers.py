"""Shared test helpers: C-oracle wrappers and error metrics."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_orc = None


def orc():
    global _orc
    if _orc is None:
        path = os.path.join(ROOT, "oracle", "liboracle.so")
        if not os.path.exists(path):
            from oracle.build_oracle import build
            build()
        _orc = ctypes.CDLL(path)
    return _orc


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _i64(*v):
    return (ctypes.c_int64 * len(v))(*v)


def c_forward(H, V, CA=None, threads=8):
    """C oracle FP64 forward (forwardProjectACC / forwardProjectGPU semantics)."""
    Hf = np.asfortranarray(H, dtype=np.float64)
    V = np.asarray(V, dtype=np.float64)
    if V.ndim == 2:
        V = V[:, :, None]
    Vf = np.asfortranarray(V)
    X, Y, nz = Vf.shape
    O = np.zeros((X, Y), order="F")
    ca = None if CA is None else np.asfortranarray(np.asarray(CA, dtype=np.int32))
    orc().orc_forward_f64(_p(Hf), _i64(*Hf.shape), None if ca is None else _p(ca), _p(Vf), ctypes.c_int64(X),
                          ctypes.c_int64(Y), ctypes.c_int64(nz), _p(O), ctypes.c_int(threads))
    return O


def c_forward_h32(H32, V, CA=None, threads=8):
    """C oracle forward with an FP32 PSF and FP64 arithmetic (no FP64 copy of H)."""
    Hf = np.asfortranarray(H32, dtype=np.float32)
    Vf = np.asfortranarray(np.asarray(V, dtype=np.float64).reshape(V.shape[0], V.shape[1], -1))
    X, Y, nz = Vf.shape
    O = np.zeros((X, Y), order="F")
    ca = None if CA is None else np.asfortranarray(np.asarray(CA, dtype=np.int32))
    orc().orc_forward_h32(_p(Hf), _i64(*Hf.shape), None if ca is None else _p(ca), _p(Vf), ctypes.c_int64(X),
                          ctypes.c_int64(Y), ctypes.c_int64(nz), _p(O), ctypes.c_int(threads))
    return O


def c_backward(H, O, CA=None, threads=8):
    Hf = np.asfortranarray(H, dtype=np.float64)
    Of = np.asfortranarray(O, dtype=np.float64)
    X, Y = Of.shape
    nz = Hf.shape[4]
    B = np.zeros((X, Y, nz), order="F")
    ca = None if CA is None else np.asfortranarray(np.asarray(CA, dtype=np.int32))
    orc().orc_backward_f64(_p(Hf), _i64(*Hf.shape), None if ca is None else _p(ca), _p(Of), ctypes.c_int64(X),
                           ctypes.c_int64(Y), ctypes.c_int64(nz), _p(B), ctypes.c_int(threads))
    return B


def c_isra(H, Htf, X0, iters, CA=None, threads=8):
    Hf = np.asfortranarray(H, dtype=np.float64)
    h = np.asfortranarray(Htf, dtype=np.float64)
    X = np.array(X0, dtype=np.float64, order="F", copy=True)
    Xs, Y, nz = X.shape
    work = np.zeros(Xs * Y + Xs * Y * nz)
    ca = None if CA is None else np.asfortranarray(np.asarray(CA, dtype=np.int32))
    orc().orc_isra_f64(_p(Hf), _i64(*Hf.shape), None if ca is None else _p(ca), _p(h), _p(X), ctypes.c_int64(Xs),
                       ctypes.c_int64(Y), ctypes.c_int64(nz), ctypes.c_int(iters), _p(work), ctypes.c_int(threads))
    return X


def rel_l2(a, ref):
    """||a - ref|| / ||ref|| (real or complex)."""
    a = np.asarray(a)
    ref = np.asarray(ref)
    cplx = np.iscomplexobj(a) or np.iscomplexobj(ref)
    a = a.astype(np.complex128 if cplx else np.float64)
    ref = ref.astype(np.complex128 if cplx else np.float64)
    return float(np.linalg.norm((a - ref).ravel()) / max(np.linalg.norm(ref.ravel()), 1e-300))


def random_psf(rng, M, n, nz, sparse_outside=None):
    """Random positive kernels; if sparse_outside = (lo, hi) zero everything outside."""
    H = rng.random((M, M, n, n, nz))
    if sparse_outside is not None:
        lo, hi = sparse_outside
        mask = np.zeros((M, M, 1, 1, 1))
        mask[lo:hi + 1, lo:hi + 1] = 1
        H = H * mask
    return np.asfortranarray(H)
