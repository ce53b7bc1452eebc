"""Seeded synthetic inputs for the light-field projection workloads.

The reference ships no PSF or data files, so every workload here is
synthetic and deterministic:

* PSF ``H`` with MATLAB layout ``[M, M, Nnum, Nnum, Nz]`` (the layout of the
  ``psf_ballistic.H`` array consumed by external/LFrecon_util/forwardProjectGPU.m:1).
  Kernel (a, b, plane z) is a 2-D Gaussian envelope of sigma ``1.5 + 0.8|z|``
  px times a lenslet-period lattice of spots placed at the lenslet centres
  seen from a source at in-lenslet offset (a, b).  The spot lattice is set
  to zero where it falls below 1e-4 of its peak, the footprint is the square
  of half-width ``min(ceil(8 sigma), floor(M/2))`` px, and each kernel is
  normalised to unit sum.  ``CAindex`` (1-based, ``Nz x 2``) is that
  footprint, the per-plane crop external/LFrecon_util/forwardProjectACC.m:22
  applies.
* Volume: seeded balls with lognormal intensities on a constant background.
* Image: forward projection scaled to a photon peak, Poisson sampled.

Arrays are numpy; ``make_psf`` also accepts ``xp=torch`` with a device to
build the multi-GB full-size PSF on the GPU.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Workload:
    """One of the five standard workloads (sizes from the workload table)."""
    name: str
    n: int                 # Nnum, pixels per lenslet side
    X: int                 # sensor rows
    Y: int                 # sensor columns
    zs: tuple              # plane z positions (px)
    M: int                 # kernel array side
    n_balls: int = 40
    r_lo: int = 3
    r_hi: int = 3
    seed: int = 0
    peak_photons: float = 1000.0
    background: float = 0.01
    iters: int = 8
    batch: int = 1
    extra: dict = field(default_factory=dict)

    @property
    def nz(self) -> int:
        return len(self.zs)


WORKLOADS = {
    # plain RL, small: 255x255 sensor (17x17 lenslets), 11 planes z=-5..5, 105 px kernels
    "rl_small": Workload("rl_small", 15, 255, 255, tuple(range(-5, 6)), 105, 40, 3, 3, seed=1),
    # plain / precision RL, full size: 1005x1005 (67x67 lenslets), 51 planes z=-25..25, 345 px kernels
    "rl_full": Workload("rl_full", 15, 1005, 1005, tuple(range(-25, 26)), 345, 3000, 3, 5, seed=2),
    # batched library: 2500 seeds in the full-size geometry
    "library": Workload("library", 15, 1005, 1005, tuple(range(-25, 26)), 345, seed=3,
                        extra=dict(n_seeds=2500, box_lo=7, box_hi=21, dz_lo=3, dz_hi=9, sigma=2.0)),
    # differentiation: 495x495 (33x33 lenslets), 31 planes, batch of 16 frames, 225 px kernels
    "grad": Workload("grad", 15, 495, 495, tuple(range(-15, 16)), 225, 40, 3, 3, seed=4, batch=16,
                     extra=dict(fd_eps=1e-3, fd_coords=64, steps=20)),
}


def plane_sigma(z: float) -> float:
    return 1.5 + 0.8 * abs(z)


def plane_radius(z: float, M: int, foot_k: float = 8.0) -> int:
    return int(min(math.ceil(foot_k * plane_sigma(z)), M // 2))


def ca_index(zs, M, foot_k: float = 8.0) -> np.ndarray:
    """1-based symmetric crop [m+1-R, m+1+R] per plane (CAindex analogue)."""
    m = M // 2
    out = np.zeros((len(zs), 2), dtype=np.int32)
    for c, z in enumerate(zs):
        R = plane_radius(z, M, foot_k)
        out[c] = (m + 1 - R, m + 1 + R)
    return out


def _spot_profile(xp, d, n, a, s):
    """Spot lattice along one axis for in-lenslet offset a: spots sit at
    d = (n-1)/2 - a (mod n), the lenslet centres relative to the source."""
    o = (n - 1) // 2 - a
    w = ((d - o + n // 2) % n) - n // 2
    return xp.exp(-(w * w) / (2.0 * s * s))


def make_psf(n: int, M: int, zs, spot_sigma: float = 1.5, foot_k: float = 8.0,
             trunc: float = 1e-4, dtype=np.float32, xp=np, device=None):
    """Synthetic PSF; returns (H, CAindex).

    numpy: H is a Fortran-ordered array of shape (M, M, n, n, Nz).
    torch: H is a contiguous tensor of shape (Nz, n, n, M, M) on ``device``,
    i.e. the same column-major memory as MATLAB's [M, M, n, n, Nz].
    """
    m = M // 2
    nz = len(zs)
    CA = ca_index(zs, M, foot_k)
    if xp is np:
        H = np.zeros((M, M, n, n, nz), dtype=dtype, order="F")
        d = np.arange(M, dtype=np.float64) - m
        S = np.stack([_spot_profile(np, d, n, a, spot_sigma) for a in range(n)])  # (n, M)
        for c, z in enumerate(zs):
            sig = plane_sigma(z)
            R = plane_radius(z, M, foot_k)
            E = np.exp(-(d * d) / (2 * sig * sig))
            E[np.abs(d) > R] = 0.0
            for b in range(n):
                for a in range(n):
                    spots = S[a][:, None] * S[b][None, :]
                    K = (E[:, None] * E[None, :]) * spots
                    K[spots < trunc] = 0.0
                    K /= K.sum()
                    H[:, :, a, b, c] = K
        return H, CA
    import torch
    dev = torch.device(device) if device is not None else torch.device("cpu")
    tdt = {np.float32: torch.float32, np.float64: torch.float64}.get(dtype, dtype)
    H = torch.empty((nz, n, n, M, M), dtype=tdt, device=dev)
    d = torch.arange(M, dtype=torch.float64, device=dev) - m
    S = torch.stack([_spot_profile(torch, d, n, a, spot_sigma) for a in range(n)])  # (n, M)
    for c, z in enumerate(zs):
        sig = plane_sigma(z)
        R = plane_radius(z, M, foot_k)
        E = torch.exp(-(d * d) / (2 * sig * sig))
        E[d.abs() > R] = 0.0
        # memory order (b, a, j, i): K[b, a, j, i] = E[i] E[j] S[a][i] S[b][j]
        spots = S[None, :, None, :] * S[:, None, :, None]            # (b, a, j, i)
        K = (E[None, None, :, None] * E[None, None, None, :]) * spots
        K[spots < trunc] = 0.0
        K /= K.sum(dim=(2, 3), keepdim=True)
        H[c] = K.to(tdt)
        del spots, K
    return H, CA


def make_volume(X: int, Y: int, nz: int, n_balls: int, r_lo: int, r_hi: int, seed: int,
                background: float = 0.01, dtype=np.float32):
    """Balls of integer radius in [r_lo, r_hi] at seeded uniform centres with
    lognormal(0, 0.5) intensities (overlaps add), plus a constant background.
    Returns a Fortran-ordered (X, Y, nz) array."""
    rng = np.random.default_rng(seed)
    cx = rng.uniform(0, X, n_balls)
    cy = rng.uniform(0, Y, n_balls)
    cz = rng.uniform(0, nz, n_balls)
    rad = rng.integers(r_lo, r_hi + 1, n_balls)
    amp = rng.lognormal(0.0, 0.5, n_balls)
    V = np.zeros((X, Y, nz), dtype=np.float64, order="F")
    for k in range(n_balls):
        r = int(rad[k])
        x0, y0, z0 = int(math.floor(cx[k])), int(math.floor(cy[k])), int(math.floor(cz[k]))
        xs = np.arange(max(x0 - r, 0), min(x0 + r + 1, X))
        ys = np.arange(max(y0 - r, 0), min(y0 + r + 1, Y))
        zz = np.arange(max(z0 - r, 0), min(z0 + r + 1, nz))
        if xs.size == 0 or ys.size == 0 or zz.size == 0:
            continue
        dx = (xs - x0)[:, None, None]
        dy = (ys - y0)[None, :, None]
        dz = (zz - z0)[None, None, :]
        inside = (dx * dx + dy * dy + dz * dz) <= r * r
        V[np.ix_(xs, ys, zz)] += amp[k] * inside
    V += background
    return np.asfortranarray(V.astype(dtype))


def poisson_image(clean, peak_photons: float, seed: int):
    """Scale to a photon peak and Poisson-sample (FP64 scale)."""
    clean = np.asarray(clean, dtype=np.float64)
    rng = np.random.default_rng(seed + 1000)
    lam = clean * (peak_photons / clean.max())
    return np.asfortranarray(rng.poisson(np.maximum(lam, 0.0)).astype(np.float64))


def make_seeds(wl: Workload, seed_offset: int = 0):
    """Library seeds: integer lateral box side in [box_lo, box_hi], depth in
    [dz_lo, dz_hi] planes, Gaussian blob (sigma px), seeded float centres.
    Returns a dict of numpy arrays (one row per seed)."""
    e = wl.extra
    rng = np.random.default_rng(wl.seed + seed_offset)
    ns = int(e["n_seeds"])
    box = rng.integers(e["box_lo"], e["box_hi"] + 1, ns)
    depth = rng.integers(e["dz_lo"], e["dz_hi"] + 1, ns)
    cx = rng.uniform(1, wl.X, ns)
    cy = rng.uniform(1, wl.Y, ns)
    cz = rng.uniform(1, wl.nz, ns)
    return dict(box=box, depth=depth, center=np.stack([cx, cy, cz], axis=1), sigma=float(e["sigma"]))
