"""Differentiation of the light-field forward model (no reference counterpart).

* ``poisson_nll_grad``: Poisson negative log-likelihood of a batch of frames
  and its gradients with respect to the volumes and the PSF, from the HIP
  adjoint kernels (sid_poisson_nll_grad_device).
* ``Project`` / ``PoissonNLL``: ``torch.autograd.Function`` wrappers over the
  same C ABI.
* ``refine``: Richardson-Lucy (Poisson EM) updates of the volumes and,
  optionally, of the PSF, with an atomic checkpoint after every step so that
  a resumed run reproduces the uninterrupted one bit for bit.

Device tensors follow the package convention: a batch of nb volumes is a
contiguous (nb, Nz, Y, X) tensor (nb MATLAB [X, Y, Nz] arrays back to back),
images (nb, Y, X), H (Nz, Nnum, Nnum, M, M).
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import os
import weakref

import numpy as np

from ._lib import check, i64, lib


def _torch():
    import torch
    return torch


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def poisson_nll_grad(psf, V, y, want_gV: bool = True, want_gH: bool = True):
    """Returns (nll, gV, gH): nll as a Python float (FP64 sum), gV like V,
    gH like H (Nz, Nnum, Nnum, M, M), or None when not requested."""
    torch = _torch()
    V = V.detach().to(torch.float32).contiguous()
    y = y.detach().to(torch.float32).contiguous()
    if V.dim() == 3:
        V, y = V.unsqueeze(0), y.unsqueeze(0)
    nb, nz, Y, X = V.shape
    gV = torch.empty_like(V) if want_gV else None
    gH = torch.empty((psf.Nz, psf.Nnum, psf.Nnum, psf.M, psf.M), dtype=torch.float32, device=V.device) \
        if want_gH else None
    nll = ctypes.c_double(0.0)
    check(lib().sid_poisson_nll_grad_device(psf._h, _ptr(V), _ptr(y), i64(X, Y, nz, nb), ctypes.byref(nll),
                                            _ptr(gV), _ptr(gH), _stream()))
    return nll.value, gV, gH


def project_vjp(psf, V, G, want_gV: bool = True, want_gH: bool = True):
    """gV = B(G) and gH = d<G, F_H(V)>/dH for upstream image gradients G."""
    torch = _torch()
    G = G.detach().to(torch.float32).contiguous()
    if G.dim() == 2:
        G = G.unsqueeze(0)
    nb, Y, X = G.shape
    if V is not None:
        V = V.detach().to(torch.float32).contiguous()
        if V.dim() == 3:
            V = V.unsqueeze(0)
    gV = torch.empty((nb, psf.Nz, Y, X), dtype=torch.float32, device=G.device) if want_gV else None
    gH = torch.empty((psf.Nz, psf.Nnum, psf.Nnum, psf.M, psf.M), dtype=torch.float32, device=G.device) \
        if want_gH else None
    check(lib().sid_project_vjp_device(psf._h, _ptr(V), _ptr(G), i64(X, Y, psf.Nz, nb), _ptr(gV), _ptr(gH),
                                       _stream()))
    return gV, gH


def backward_ones(psf, X: int, Y: int):
    """FP64 B(1) = backwardProjectGPU_new(H, ones(X, Y)) as a (Nz, Y, X) float64
    tensor: exact sums of each kernel over its part that lands on the sensor."""
    torch = _torch()
    out = torch.empty((psf.Nz, Y, X), dtype=torch.float64, device="cuda")
    check(lib().sid_backward_ones_device(psf._h, int(X), int(Y), _ptr(out), _stream()))
    return out


def _sync_psf(psf, H):
    """Keep the handle's H equal to the tensor H (rebuild weights on change).

    The handle is current when H is the very tensor it was last built from
    (held by a weak reference, so a new tensor that reuses a freed block's
    address never matches) and H has not been modified in place since (its
    version counter).  Writes through ``H.data`` bypass the version counter;
    call ``psf.update(H)`` after them."""
    ref = getattr(psf, "_h_ref", None)
    if ref is None or ref() is not H or psf._h_version != H._version:
        psf.update(H)
        psf._h_ref = weakref.ref(H)
        psf._h_version = H._version


def _make_functions():
    torch = _torch()

    class Project(torch.autograd.Function):
        """O = forward(H, V) for a batch (nb, Nz, Y, X) -> (nb, Y, X)."""

        @staticmethod
        def forward(ctx, V, H, psf):
            _sync_psf(psf, H)
            outs = [psf.forward(V[f]) for f in range(V.shape[0])]
            ctx.save_for_backward(V)
            ctx.psf = psf
            return torch.stack(outs)

        @staticmethod
        def backward(ctx, gO):
            (V,) = ctx.saved_tensors
            gV, gH = project_vjp(ctx.psf, V, gO, ctx.needs_input_grad[0], ctx.needs_input_grad[1])
            return gV, gH, None

    class PoissonNLL(torch.autograd.Function):
        """sum_f sum_p F_f - y_f log F_f as a float64 scalar."""

        @staticmethod
        def forward(ctx, V, H, y, psf):
            _sync_psf(psf, H)
            nll, gV, gH = poisson_nll_grad(psf, V, y, True, True)
            ctx.save_for_backward(gV, gH)
            return torch.tensor(nll, dtype=torch.float64, device=V.device)

        @staticmethod
        def backward(ctx, g):
            gV, gH = ctx.saved_tensors
            g32 = g.to(torch.float32)
            return g32 * gV, g32 * gH, None, None

    return Project, PoissonNLL


_FUNCS = None


def functions():
    """(Project, PoissonNLL) autograd Functions (built on first use)."""
    global _FUNCS
    if _FUNCS is None:
        _FUNCS = _make_functions()
    return _FUNCS


# ---------------------------------------------------------------------------
# Refinement loop with per-iteration checkpointing
# ---------------------------------------------------------------------------

def _config_hash(cfg: dict) -> str:
    return hashlib.sha256(json.dumps(cfg, sort_keys=True).encode()).hexdigest()[:16]


def save_checkpoint(path: str, step: int, V, H, history, cfg: dict):
    """Atomic write (tmp file + rename) of the refinement state."""
    from safetensors.torch import save_file
    torch = _torch()
    tensors = {"V": V.detach().contiguous(), "H": H.detach().contiguous(),
               "nll": torch.tensor(history, dtype=torch.float64)}
    meta = {"step": str(step), "config": json.dumps(cfg, sort_keys=True), "config_hash": _config_hash(cfg)}
    tmp = path + ".tmp"
    save_file({k: v.cpu() for k, v in tensors.items()}, tmp, metadata=meta)
    os.replace(tmp, path)


def load_checkpoint(path: str, device):
    from safetensors import safe_open
    with safe_open(path, framework="pt") as f:
        meta = f.metadata()
        V = f.get_tensor("V").to(device)
        H = f.get_tensor("H").to(device)
        hist = f.get_tensor("nll").tolist()
    return int(meta["step"]), V, H, hist, json.loads(meta["config"])


def refine(psf, H, V0, y, steps: int, checkpoint: str | None = None, refine_psf: bool = True,
           resume: bool = True, stop_after: int | None = None, checkpoint_every: int = 1):
    """Poisson-EM (Richardson-Lucy) refinement of volumes and PSF.

    Each step is two EM half-steps, each of which never increases the NLL:
      V <- V * (1 - gV / B(1))   with gV = dNLL/dV at (V, H)   (= V * B(y/F) / B(1))
      H <- H * (1 - gH / G1)     with gH = dNLL/dH at (V_new, H) and
                                 G1 = d(sum F)/dH at V_new      (if refine_psf)
    B(1) is the back-projected normalisation, exact in FP64 (backward_ones;
    recomputed when H changes); the V ratio is formed in FP64.
    The NLL recorded for a step is the one at the start of the step.  After
    each step (every ``checkpoint_every`` steps, and after the last or the
    ``stop_after`` step) the state is checkpointed atomically to ``checkpoint``;
    ``stop_after`` ends the loop early (to test resumption).
    Returns (V, H, nll_history).
    """
    torch = _torch()
    cfg = {"steps": int(steps), "refine_psf": bool(refine_psf), "shape": list(V0.shape), "psf": list(H.shape)}
    V = V0.detach().to(torch.float32).contiguous().clone()
    H = H.detach().to(torch.float32).contiguous().clone()
    hist = []
    start = 0
    if checkpoint and resume and os.path.exists(checkpoint):
        start, Vc, Hc, hist, cfg_saved = load_checkpoint(checkpoint, V.device)
        if _config_hash(cfg_saved) != _config_hash(cfg):
            raise ValueError("checkpoint was written by a different configuration")
        V.copy_(Vc)
        H.copy_(Hc)
    psf.update(H)
    psf.set_method("direct")   # the PSF gradient needs the direct projector; keep one plan
    nb, nz, Y, X = V.shape
    ones = torch.ones((nb, Y, X), dtype=torch.float32, device=V.device)
    B1 = backward_ones(psf, X, Y).unsqueeze(0)
    for step in range(start, steps):
        nll, gV, _ = poisson_nll_grad(psf, V, y, True, False)
        hist.append(nll)
        with torch.no_grad():
            ratio = torch.where(B1 > 0, gV.double() / B1, torch.zeros_like(B1)).to(torch.float32)
            V = torch.clamp(V * (1.0 - ratio), min=0.0).contiguous()
        if refine_psf:
            _, _, gH = poisson_nll_grad(psf, V, y, False, True)
            _, G1 = project_vjp(psf, V, ones, False, True)
            with torch.no_grad():
                rh = torch.where(G1 > 0, gH / G1, torch.zeros_like(gH))
                H = torch.clamp(H * (1.0 - rh), min=0.0).contiguous()
            psf.update(H)
            B1 = backward_ones(psf, X, Y).unsqueeze(0)
        stop = stop_after is not None and step + 1 >= stop_after
        if checkpoint and ((step + 1) % max(1, int(checkpoint_every)) == 0 or step + 1 == steps or stop):
            save_checkpoint(checkpoint, step + 1, V, H, hist, cfg)
        if stop:
            break
    return V, H, hist
