"""MI355X light-field projection path of SID (vazirilab/SID), host-side mirror.

The functions below keep the names, argument order and meaning of the
reference MATLAB functions they replace and run on libsidproj.so (HIP,
gfx950).  There is no CPU fallback: without the library or a gfx950 device
every call raises.

    forwardProjectGPU(H, realspace)            external/LFrecon_util/forwardProjectGPU.m:1
    backwardProjectGPU_new(H, projection)      external/LFrecon_util/backwardProjectGPU_new.m:1
    forwardProjectACC(H, realspace, CAindex)   external/LFrecon_util/forwardProjectACC.m:1
    backwardProjectACC_new(H, proj, CAindex)   external/LFrecon_util/backwardProjectACC_new.m:1
    deconvRL(forwardFUN, backwardFUN, Htf, maxIter, Xguess, opts)
                                               external/LFrecon_util/deconvRL.m:1
    LS_deconv(forwardFUN, backwardFUN, df, opts)
                                               external/LFrecon_util/LS_deconv.m:1

``H`` may be a numpy array indexed like MATLAB (shape (M, M, Nnum, Nnum, Nz)),
a device tensor of shape (Nz, Nnum, Nnum, M, M) (the same column-major
memory), or a ``PSF`` handle, which keeps H and the projector weights
resident on the device across calls.  Volumes and images are numpy arrays
indexed like MATLAB ((X, Y, nz) and (X, Y)) or device tensors in the reversed
shape ((nz, Y, X), (Y, X)); results come back in the same kind.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import PREC, METHOD, SidError, check, i64, lib

__all__ = ["PSF", "forwardProjectGPU", "backwardProjectGPU_new", "forwardProjectACC",
           "backwardProjectACC_new", "deconvRL", "LS_deconv", "convn_same", "SidError", "device_count", "selftest"]


def device_count() -> int:
    return int(lib().sid_device_count())


def selftest() -> None:
    check(lib().sid_selftest())


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def _torch():
    import torch
    return torch


def _stream_ptr():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class PSF:
    """Device-resident PSF handle (``sid_psf``).

    ``CAindex`` (Nz x 2, 1-based) selects the cropped CPU-path semantics of
    forwardProjectACC/backwardProjectACC_new; ``None`` the full-kernel GPU
    path.  ``precision`` is 'fp32', 'bf16x2' or 'bf16x3'.
    """

    def __init__(self, H, CAindex=None, precision: str = "fp32", method: str = "auto", full_support: bool = False):
        self._h = ctypes.c_void_p()
        flags = 2 if full_support else 0      # SID_FLAG_FULL_SUPPORT
        ca = None
        if CAindex is not None:
            ca_arr = np.asfortranarray(np.asarray(CAindex, dtype=np.int32))
            self._ca = ca_arr
            ca = ca_arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        if _is_torch(H):
            torch = _torch()
            if H.dim() != 5 or not H.is_cuda:
                raise ValueError("device H must be a 5-D cuda tensor of shape (Nz, Nnum, Nnum, M, M)")
            Hc = H.detach().to(torch.float32).contiguous()
            nz, n, n2, M, M2 = Hc.shape
            dims = i64(M2, M, n2, n, nz)
            torch.cuda.synchronize()
            check(lib().sid_psf_create_ex(ctypes.c_void_p(Hc.data_ptr()), dims, ca, flags | 1, ctypes.byref(self._h)))
        else:
            Ha = np.asfortranarray(np.asarray(H, dtype=np.float32))
            if Ha.ndim != 5:
                raise ValueError("H must have 5 dimensions [M, M, Nnum, Nnum, Nz]")
            dims = i64(*Ha.shape)
            check(lib().sid_psf_create_ex(Ha.ctypes.data_as(ctypes.c_void_p), dims, ca, flags, ctypes.byref(self._h)))
        self.dims = tuple(int(d) for d in dims)
        self.M, _, self.Nnum, _, self.Nz = self.dims
        self.precision = "fp32"
        self.kernel = None
        self.set_precision(precision)
        if method != "auto":
            self.set_method(method)

    @classmethod
    def from_file(cls, path: str, precision: str = "fp32", method: str = "auto", use_caindex: bool = False):
        """Handle for the PSF MAT-file at path (``H``, ``Nnum``, ``CAindex``;
        -v7.3 HDF5 or older MAT formats, see sid_amd.psf_io).  The GPU path of
        reconstruct_S.m:155 takes H only; ``use_caindex`` applies the file's
        CAindex crops as the CPU path does (reconstruct_S.m:118-121)."""
        from .psf_io import load_psf
        H, _, ca = load_psf(path)
        if use_caindex and ca is None:
            raise ValueError(f"{path}: use_caindex asked for but the file has no CAindex")
        return cls(H, ca if use_caindex else None, precision=precision, method=method)

    # -- configuration -----------------------------------------------------
    def update(self, H):
        """Replace the values of H (same shape) and rebuild the weights."""
        if _is_torch(H):
            torch = _torch()
            Hc = H.detach().to(torch.float32).contiguous()
            torch.cuda.synchronize()
            check(lib().sid_psf_update(self._h, ctypes.c_void_p(Hc.data_ptr()), 1))
        else:
            Ha = np.asfortranarray(np.asarray(H, dtype=np.float32))
            check(lib().sid_psf_update(self._h, Ha.ctypes.data_as(ctypes.c_void_p), 0))

    def set_precision(self, precision: str):
        check(lib().sid_psf_set_precision(self._h, PREC[precision]))
        self.precision = precision

    def set_method(self, method: str):
        check(lib().sid_psf_set_method(self._h, METHOD[method]))

    def set_overlap(self, on: bool):
        """Use the handle's second stream beside the caller's (default) or not;
        the sensor plan is rebuilt by the next projection."""
        check(lib().sid_psf_set_overlap(self._h, 1 if on else 0))

    def set_plane_methods(self, methods=None):
        """Pin planes to "direct" / "fft" ("auto" leaves a plane to the
        autotuner); one entry per plane of H, None clears every pin."""
        if methods is None:
            check(lib().sid_psf_set_plane_methods(self._h, None, 0))
            return
        arr = (ctypes.c_int32 * len(methods))(*[METHOD[m] for m in methods])
        check(lib().sid_psf_set_plane_methods(self._h, arr, len(methods)))

    def info(self) -> dict:
        inf = _lib.PsfInfo()
        check(lib().sid_psf_get_info(self._h, ctypes.byref(inf)))
        names = {0: "empty", 1: "direct", 2: "fft"}
        return dict(M=inf.M, Nnum=inf.Nnum, Nz=inf.Nz, taps_max=inf.taps_max, weight_bytes=inf.weight_bytes,
                    support=[(inf.support_lo[c], inf.support_hi[c]) for c in range(inf.Nz)],
                    plane_method=[names.get(inf.plane_method[c], "?") for c in range(inf.Nz)],
                    autotune_ms=list(inf.autotune_ms), spectra_bytes=inf.spectra_bytes,
                    fft_tile=list(inf.fft_tile),
                    fft_groups=[(inf.fft_group_planes[g], inf.fft_group_T[g]) for g in range(min(inf.fft_groups, 8))],
                    fft_gemm_bytes=list(inf.fft_gemm_bytes), fft_gemm_flops=list(inf.fft_gemm_flops))

    def timing(self) -> list:
        buf = (ctypes.c_double * 12)()
        check(lib().sid_last_timing(self._h, buf))
        return list(buf)

    def plan_stats(self) -> dict:
        buf = (ctypes.c_int64 * 8)()
        check(lib().sid_plan_stats(self._h, buf))
        keys = ["macs_fwd", "macs_bwd", "items_fwd", "items_bwd", "slabs_fwd", "slabs_bwd", "rows_padded",
                "cols_padded"]
        return dict(zip(keys, list(buf)))

    def plane_macs(self):
        """MFMA multiply-adds per plane of the current plan's forward and
        backward GEMM launches (executed K rows; 0 for FFT or empty planes)."""
        f = (ctypes.c_int64 * self.Nz)()
        b = (ctypes.c_int64 * self.Nz)()
        check(lib().sid_plan_plane_macs(self._h, f, b, self.Nz))
        return list(f), list(b)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().sid_psf_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- projectors ---------------------------------------------------------
    def forward(self, realspace):
        """forwardProjectGPU(H, realspace): (X, Y, nz) -> (X, Y) float32."""
        if _is_torch(realspace):
            torch = _torch()
            V = realspace.detach().to(torch.float32).contiguous()
            if V.dim() == 2:
                V = V.unsqueeze(0)
            nz, Y, X = V.shape
            out = torch.empty((Y, X), dtype=torch.float32, device=V.device)
            check(lib().sid_forward_device(self._h, ctypes.c_void_p(V.data_ptr()), i64(X, Y, nz),
                                           ctypes.c_void_p(out.data_ptr()), _stream_ptr()))
            return out
        V = np.asarray(realspace)
        if V.ndim == 2:
            V = V[:, :, None]
        V = np.asfortranarray(V, dtype=np.float32)
        X, Y, nz = V.shape
        out = np.zeros((X, Y), dtype=np.float32, order="F")
        check(lib().sid_forward(self._h, V.ctypes.data_as(ctypes.c_void_p), i64(X, Y, nz),
                                out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def backward(self, projection):
        """backwardProjectGPU_new(H, projection): (X, Y) -> (X, Y, Nz) float32."""
        if _is_torch(projection):
            torch = _torch()
            O = projection.detach().to(torch.float32).contiguous()
            Y, X = O.shape
            out = torch.empty((self.Nz, Y, X), dtype=torch.float32, device=O.device)
            check(lib().sid_backward_device(self._h, ctypes.c_void_p(O.data_ptr()), i64(X, Y),
                                            ctypes.c_void_p(out.data_ptr()), _stream_ptr()))
            return out
        O = np.asfortranarray(np.asarray(projection), dtype=np.float32)
        X, Y = O.shape
        out = np.zeros((X, Y, self.Nz), dtype=np.float32, order="F")
        check(lib().sid_backward(self._h, O.ctypes.data_as(ctypes.c_void_p), i64(X, Y),
                                 out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def deconv_rl(self, Htf, maxIter: int, Xguess, history: bool = False):
        """deconvRL basic (lambda = 0) with this PSF's projectors, all on device."""
        if tuple(Xguess.shape) != tuple(Htf.shape):
            raise ValueError("deconv_rl: Xguess has shape %s but Htf has %s" % (tuple(Xguess.shape), tuple(Htf.shape)))
        if _is_torch(Htf):
            torch = _torch()
            h = Htf.detach().to(torch.float32).contiguous()
            X = Xguess.detach().to(torch.float32).contiguous().clone()
            nz, Y, Xs = h.shape
            hist = torch.empty((maxIter, nz, Y, Xs), dtype=torch.float32, device=h.device) if history else None
            check(lib().sid_deconv_rl_device(self._h, ctypes.c_void_p(h.data_ptr()), ctypes.c_void_p(X.data_ptr()),
                                             i64(Xs, Y, nz), int(maxIter),
                                             ctypes.c_void_p(hist.data_ptr()) if history else None, _stream_ptr()))
            return (X, hist) if history else X
        h = np.asfortranarray(np.asarray(Htf), dtype=np.float32)
        X = np.array(np.asarray(Xguess), dtype=np.float32, order="F", copy=True)
        Xs, Y, nz = h.shape
        hist = np.zeros((Xs * Y * nz * max(int(maxIter), 0),), dtype=np.float32) if history else None
        check(lib().sid_deconv_rl(self._h, h.ctypes.data_as(ctypes.c_void_p), X.ctypes.data_as(ctypes.c_void_p),
                                  i64(Xs, Y, nz), int(maxIter),
                                  hist.ctypes.data_as(ctypes.c_void_p) if history else None))
        if history:
            # Xguess_history(i, :, :, :) of deconvRL.m:10,91
            hist = hist.reshape((Xs, Y, nz, int(maxIter)), order="F").transpose(3, 0, 1, 2)
            return X, hist
        return X

    def ls_deconv(self, df, opts=None):
        """LS_deconv (projected gradient descent, exact line search) with this
        PSF's projectors, all on device.  ``opts`` keys as in LS_deconv.m:25-43."""
        o = _ls_opts(opts)
        if _is_torch(df):
            torch = _torch()
            h = df.detach().to(torch.float32).contiguous()
            X = torch.empty_like(h)
            nz, Y, Xs = h.shape
            check(lib().sid_ls_deconv_device(self._h, ctypes.c_void_p(h.data_ptr()), ctypes.c_void_p(X.data_ptr()),
                                             i64(Xs, Y, nz), ctypes.byref(o), _stream_ptr()))
            return X
        h = np.asfortranarray(np.asarray(df), dtype=np.float32)
        X = np.zeros(h.shape, dtype=np.float32, order="F")
        Xs, Y, nz = h.shape
        check(lib().sid_ls_deconv(self._h, h.ctypes.data_as(ctypes.c_void_p), X.ctypes.data_as(ctypes.c_void_p),
                                  i64(Xs, Y, nz), ctypes.byref(o)))
        return X

    def set_kernel(self, kernel):
        """Neuron kernel of reconstruct_S.m:122-128,184-190: afterwards every
        projection is F(convn(V, k, 'same')) / convn(B(O), k, 'same').
        ``kernel`` is indexed like MATLAB, shape (kx, ky, kz) (a 2-D array is
        one plane); None removes it."""
        if kernel is None:
            check(lib().sid_psf_set_kernel(self._h, None, None))
            self.kernel = None
            return
        k = np.asarray(kernel, dtype=np.float32)
        if k.ndim == 2:
            k = k[:, :, None]
        k = np.asfortranarray(k)
        check(lib().sid_psf_set_kernel(self._h, k.ctypes.data_as(ctypes.c_void_p), i64(*k.shape)))
        self.kernel = k

    def forward_fun(self):
        f = lambda V: self.forward(V)
        f._sid_psf = self
        f._sid_dir = "forward"
        return f

    def backward_fun(self):
        f = lambda O: self.backward(O)
        f._sid_psf = self
        f._sid_dir = "backward"
        return f


def convn_same(V, k, precision: str = "fp32"):
    """convn(V, k, 'same') on the device (MATLAB window: offset
    floor(size(k)/2)); precision 'fp32' (FP32 arithmetic) or 'bf16x2' /
    'bf16x3' (BF16-split MFMA, as a handle in that mode computes it).  numpy
    arrays indexed like MATLAB ((X, Y, Z) and (kx, ky, kz)), or device tensors
    in the reversed shape ((Z, Y, X), (kz, ky, kx)); results come back in the
    same kind."""
    prec = PREC[precision]
    if _is_torch(V):
        torch = _torch()
        v = V.detach().to(torch.float32).contiguous()
        kk = k.detach().to(device=v.device, dtype=torch.float32).contiguous() if _is_torch(k) else \
            torch.from_numpy(np.ascontiguousarray(np.asarray(k, dtype=np.float32).T)).to(v.device)
        v3 = v if v.dim() == 3 else v[None]
        k3 = kk if kk.dim() == 3 else kk[None]
        out = torch.empty_like(v3)
        check(lib().sid_convn_same_device(ctypes.c_void_p(v3.data_ptr()), i64(*reversed(v3.shape)),
                                          ctypes.c_void_p(k3.data_ptr()), i64(*reversed(k3.shape)), prec,
                                          ctypes.c_void_p(out.data_ptr()), _stream_ptr()))
        return out if v.dim() == 3 else out[0]
    v = np.asarray(V, dtype=np.float32)
    squeeze = v.ndim == 2
    if squeeze:
        v = v[:, :, None]
    kk = np.asarray(k, dtype=np.float32)
    if kk.ndim == 2:
        kk = kk[:, :, None]
    v = np.asfortranarray(v)
    kk = np.asfortranarray(kk)
    out = np.zeros(v.shape, dtype=np.float32, order="F")
    check(lib().sid_convn_same(v.ctypes.data_as(ctypes.c_void_p), i64(*v.shape), kk.ctypes.data_as(ctypes.c_void_p),
                               i64(*kk.shape), prec, out.ctypes.data_as(ctypes.c_void_p)))
    return out[:, :, 0] if squeeze else out


def _ls_opts(opts) -> "_lib.LsOpts":
    """Defaults of LS_deconv.m:25-43; lamb_TV_L2 needs all three entries, as
    the filter construction at LS_deconv.m:54-57 indexes (1)..(3)."""
    opts = dict(opts or {})
    o = _lib.LsOpts()
    o.max_iter = int(opts.get("max_iter", 8))
    o.lamb_TV_L1 = float(opts.get("lamb_TV_L1", 0.0))
    tv = np.asarray(opts.get("lamb_TV_L2", (0.0, 0.0, 0.0)), dtype=np.float64).reshape(-1)
    if tv.size == 1 and tv[0] == 0:
        tv = np.zeros(3)   # a scalar 0 (the config default) leaves the TV term off, as in MATLAB
    if tv.size != 3:
        raise ValueError("lamb_TV_L2 needs one multiplier per volume dimension")
    for i in range(3):
        o.lamb_TV_L2[i] = float(tv[i])
    o.lamb_L1 = float(opts.get("lamb_L1", 0.0))
    o.lamb_L2 = float(opts.get("lamb_L2", 0.0))
    return o


def _as_psf(H, CAindex=None) -> PSF:
    if isinstance(H, PSF):
        return H
    return PSF(H, CAindex)


def forwardProjectGPU(H, realspace):
    """external/LFrecon_util/forwardProjectGPU.m:1 — full M x M kernels."""
    return _as_psf(H).forward(realspace)


def backwardProjectGPU_new(H, projection):
    """external/LFrecon_util/backwardProjectGPU_new.m:1 — full M x M kernels."""
    return _as_psf(H).backward(projection)


def forwardProjectACC(H, realspace, CAindex):
    """external/LFrecon_util/forwardProjectACC.m:1 — kernels cropped to CAindex."""
    return (H if isinstance(H, PSF) else PSF(H, CAindex)).forward(realspace)


def backwardProjectACC_new(H, projection, CAindex):
    """external/LFrecon_util/backwardProjectACC_new.m:1 — kernels cropped to CAindex."""
    return (H if isinstance(H, PSF) else PSF(H, CAindex)).backward(projection)


def deconvRL(forwardFUN, backwardFUN, Htf, maxIter, Xguess, opts=None):
    """external/LFrecon_util/deconvRL.m:1 (mode 'basic', lambda = 0).

    ``forwardFUN``/``backwardFUN`` must be the projector handles of one
    ``PSF`` (``psf.forward_fun()``, ``psf.backward_fun()``); the whole loop
    then runs on the device.  Returns Xguess, or (Xguess, Xguess_history) when
    ``opts`` has ``history=True`` (deconvRL.m:9-11 uses nargout for that).
    """
    opts = dict(opts or {})
    mode = opts.get("mode", "basic")
    lam = opts.get("lambda", 0)
    if mode != "basic" or np.any(np.asarray(lam) != 0):
        raise NotImplementedError("only mode 'basic' with lambda = 0 is implemented on the device")
    psf_f = getattr(forwardFUN, "_sid_psf", None)
    psf_b = getattr(backwardFUN, "_sid_psf", None)
    if psf_f is None or psf_f is not psf_b:
        raise TypeError("forwardFUN and backwardFUN must be PSF.forward_fun()/backward_fun() of the same PSF")
    return psf_f.deconv_rl(Htf, int(maxIter), Xguess, history=bool(opts.get("history", False)))


def LS_deconv(forwardFUN, backwardFUN, df, opts=None):
    """external/LFrecon_util/LS_deconv.m:1 — projected gradient descent with
    exact line search for the regularised non-negative least-squares problem
    argmin_{x>=0} ||y - A x||^2 + TV_L2 + lamb_L2 ||x||^2 + lamb_L1 ||x||_1,
    started at ``df`` = A'(y).  ``forwardFUN``/``backwardFUN`` must be the
    projector handles of one ``PSF``; ``opts`` keys: max_iter, lamb_TV_L1,
    lamb_TV_L2 (3 entries), lamb_L1, lamb_L2.
    """
    psf_f = getattr(forwardFUN, "_sid_psf", None)
    psf_b = getattr(backwardFUN, "_sid_psf", None)
    if psf_f is None or psf_f is not psf_b:
        raise TypeError("forwardFUN and backwardFUN must be PSF.forward_fun()/backward_fun() of the same PSF")
    return psf_f.ls_deconv(df, opts)
