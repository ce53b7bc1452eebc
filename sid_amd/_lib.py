"""ctypes bindings of libsidproj.so (declared in include/sid_proj.h).

Loading fails loudly: there is no CPU or PyTorch fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsidproj.so")

SID_OK = 0
STATUS_NAMES = {0: "SID_OK", 1: "SID_ERR_ARG", 2: "SID_ERR_DIMS", 3: "SID_ERR_NZ", 4: "SID_ERR_OOM",
                5: "SID_ERR_HIP", 6: "SID_ERR_STATE", 7: "SID_ERR_NODEV"}
PREC = {"fp32": 0, "bf16x2": 1, "bf16x3": 2, "fp64": 3}
METHOD = {"auto": 0, "direct": 1, "fft": 2}

# Every exported symbol with (restype, argtypes); tests check this list
# against include/sid_proj.h.
_P = ctypes.c_void_p
_I64P = ctypes.POINTER(ctypes.c_int64)
_I32P = ctypes.POINTER(ctypes.c_int32)
_I = ctypes.c_int
SIGNATURES = {
    "sid_last_error": (ctypes.c_char_p, []),
    "sid_version": (_I, []),
    "sid_device_count": (_I, []),
    "sid_psf_create": (_I, [_P, _I64P, _I32P, ctypes.POINTER(_P)]),
    "sid_psf_create_device": (_I, [_P, _I64P, _I32P, ctypes.POINTER(_P)]),
    "sid_psf_destroy": (_I, [_P]),
    "sid_psf_is_live": (_I, [_P]),
    "sid_psf_create_ex": (_I, [_P, _I64P, _I32P, _I, ctypes.POINTER(_P)]),
    "sid_psf_update": (_I, [_P, _P, _I]),
    "sid_poisson_nll_grad_device": (_I, [_P, _P, _P, _I64P, ctypes.POINTER(ctypes.c_double), _P, _P, _P]),
    "sid_project_vjp_device": (_I, [_P, _P, _P, _I64P, _P, _P, _P]),
    "sid_backward_ones_device": (_I, [_P, ctypes.c_int64, ctypes.c_int64, _P, _P]),
    "sid_psf_set_precision": (_I, [_P, _I]),
    "sid_psf_set_method": (_I, [_P, _I]),
    "sid_psf_set_overlap": (_I, [_P, _I]),
    "sid_psf_set_plane_methods": (_I, [_P, _I32P, ctypes.c_int64]),
    "sid_psf_get_info": (_I, [_P, _P]),
    "sid_forward": (_I, [_P, _P, _I64P, _P]),
    "sid_backward": (_I, [_P, _P, _I64P, _P]),
    "sid_forward_device": (_I, [_P, _P, _I64P, _P, _P]),
    "sid_backward_device": (_I, [_P, _P, _I64P, _P, _P]),
    "sid_deconv_rl": (_I, [_P, _P, _P, _I64P, _I, _P]),
    "sid_deconv_rl_device": (_I, [_P, _P, _P, _I64P, _I, _P, _P]),
    "sid_library_rects": (_I, [_P, ctypes.c_int64, _I64P, _I64P, _I64P, _I64P]),
    "sid_library_forward_device": (_I, [_P, ctypes.c_int64, _I64P, _P, _I64P, _I64P, _P, _I, _P]),
    "sid_library_stats": (_I, [_P, _I64P]),
    "sid_library_forward": (_I, [_P, ctypes.c_int64, _I64P, _P, _I64P, _I64P, _P]),
    "sid_patched_rect": (_I, [ctypes.c_int64, ctypes.c_int64, _I64P, _I64P, _I64P]),
    "sid_library_kernel_boxes": (_I, [ctypes.c_int64, _I64P, _I64P, _I64P, _I64P, _I64P]),
    "sid_library_convn_seeds_device": (_I, [ctypes.c_int64, _I64P, _P, _I64P, _I64P, _P, _I64P, _P, _P]),
    "sid_library_convn_seeds": (_I, [ctypes.c_int64, _I64P, _P, _I64P, _I64P, _P, _I64P, _P]),
    "sid_library_nsf_bound": (_I, [_P, ctypes.c_int64, _P, ctypes.c_double, _P, _I64P, _P, _I64P, _I64P]),
    "sid_library_nsf": (_I, [_P, ctypes.c_int64, _P, ctypes.c_double, _P, _I64P, _P, _I64P, _I64P, _I64P, _P,
                             ctypes.c_int64]),
    "sid_psf_set_kernel": (_I, [_P, _P, _I64P]),
    "sid_convn_same_device": (_I, [_P, _I64P, _P, _I64P, _I, _P, _P]),
    "sid_convn_same": (_I, [_P, _I64P, _P, _I64P, _I, _P]),
    "sid_std_image_device": (_I, [_P, ctypes.c_int64, ctypes.c_int64, _I, _I, _P, _P, _P, _P, _P]),
    "sid_std_image": (_I, [_P, ctypes.c_int64, ctypes.c_int64, _I, _I, _P, _P, _P, _P]),
    "sid_std_image_device_noalloc": (_I, [_P, ctypes.c_int64, ctypes.c_int64, _I, _I, _P, _P, _P, _P, _P]),
    "sid_ls_deconv": (_I, [_P, _P, _P, _I64P, _P]),
    "sid_ls_deconv_device": (_I, [_P, _P, _P, _I64P, _P, _P]),
    "sid_last_timing": (_I, [_P, ctypes.POINTER(ctypes.c_double)]),
    "sid_plan_stats": (_I, [_P, _I64P]),
    "sid_plan_plane_macs": (_I, [_P, _I64P, _I64P, ctypes.c_int64]),
    "sid_selftest": (_I, []),
    "sid_fft_device": (_I, [_I, ctypes.c_int64, _P, _I, _P]),
}


class SidError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class PsfInfo(ctypes.Structure):
    _fields_ = [("M", ctypes.c_int64), ("Nnum", ctypes.c_int64), ("Nz", ctypes.c_int64),
                ("taps_max", ctypes.c_int64), ("weight_bytes", ctypes.c_int64),
                ("support_lo", ctypes.c_int64 * 512), ("support_hi", ctypes.c_int64 * 512),
                ("plane_method", ctypes.c_int32 * 512), ("autotune_ms", ctypes.c_double * 4),
                ("spectra_bytes", ctypes.c_int64),
                ("fft_tile", ctypes.c_int32 * 4),
                ("fft_groups", ctypes.c_int32),
                ("fft_group_planes", ctypes.c_int32 * 8),
                ("fft_group_T", ctypes.c_int32 * 8),
                ("fft_gemm_bytes", ctypes.c_int64 * 2),
                ("fft_gemm_flops", ctypes.c_double * 2)]


class LsOpts(ctypes.Structure):
    _fields_ = [("max_iter", ctypes.c_int), ("lamb_TV_L1", ctypes.c_double), ("lamb_TV_L2", ctypes.c_double * 3),
                ("lamb_L1", ctypes.c_double), ("lamb_L2", ctypes.c_double)]


_lib = None


def lib():
    """Load libsidproj.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with "
                              "`python -c 'import __graft_entry__ as g; g.build()'`")
        # One HIP runtime per process: when PyTorch is installed, load its
        # libamdhip64 first so this library binds to the same runtime
        # (two runtimes in one process cannot share the device).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(code: int):
    if code != SID_OK:
        raise SidError(code, lib().sid_last_error().decode(errors="replace"))


def i64(*vals):
    return (ctypes.c_int64 * len(vals))(*[int(v) for v in vals])
