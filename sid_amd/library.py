"""Library generation around the batched HIP projector.

``generate_LFM_library_CPU`` builds the normalised ball / kernel footprints of
src/generate_LFM_library_CPU.m on the same batched projection.
``placement`` restates the integer seed placement of
src/generate_LFM_library_GPU.m:68-88 (MATLAB ``round`` is half away from
zero); ``patched_rect`` calls the library's own rectangle rule of
src/project_forward_patched.m:8-20,54-59; ``Library`` projects thousands of
seed sub-volumes in one persistent, hipGraph-captured launch
(sid_library_forward_device).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from ._lib import check, i64, lib


def matlab_round(x: float) -> int:
    return int(math.copysign(math.floor(abs(x) + 0.5), x))


def placement(center, neur_rad: int, dz: int, image_size, nz: int):
    """Sub-volume box {x0, x1, y0, y1, z0, z1} (1-based inclusive) of a seed:
    x, y, z = round(centre); z in [max(z - dz, 1), min(z + dz, Nz)] with
    dz = ceil(neur_rad / axial) in the reference; lateral box
    [max(x - r, 1), min(x + r, X)] x [max(y - r, 1), min(y + r, Y)]."""
    x, y, z = (matlab_round(float(c)) for c in center)
    return (max(x - neur_rad, 1), min(x + neur_rad, image_size[0]),
            max(y - neur_rad, 1), min(y + neur_rad, image_size[1]),
            max(z - dz, 1), min(z + dz, nz))


def patched_rect(n: int, M: int, sensor, box):
    rect = (ctypes.c_int64 * 4)()
    check(lib().sid_patched_rect(int(n), int(M), i64(*sensor), i64(*box), rect))
    return tuple(rect)


def gaussian_blob(box, center, sigma: float):
    """FP32 Gaussian blob on the box (column-major), centred at the float
    centre (1-based coordinates)."""
    x = np.arange(box[0], box[1] + 1, dtype=np.float64) - center[0]
    y = np.arange(box[2], box[3] + 1, dtype=np.float64) - center[1]
    z = np.arange(box[4], box[5] + 1, dtype=np.float64) - center[2]
    g = (np.exp(-x * x / (2 * sigma * sigma))[:, None, None] * np.exp(-y * y / (2 * sigma * sigma))[None, :, None]
         * np.exp(-z * z / (2 * sigma * sigma))[None, None, :])
    return np.asfortranarray(g.astype(np.float32))


class Library:
    """Batched projection of seed sub-volumes with one PSF handle."""

    def __init__(self, psf, sensor):
        self.psf = psf
        self.sensor = (int(sensor[0]), int(sensor[1]))

    def rects(self, boxes):
        boxes = np.ascontiguousarray(np.asarray(boxes, dtype=np.int64).reshape(-1, 6))
        ns = boxes.shape[0]
        rects = np.zeros((ns, 4), dtype=np.int64)
        offs = np.zeros(ns + 1, dtype=np.int64)
        P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        check(lib().sid_library_rects(self.psf._h, ns, P(boxes), i64(*self.sensor), P(rects), P(offs)))
        return rects, offs

    def prepare(self, boxes, values):
        """Upload the seeds' values; returns (vals_dev, val_offsets, out_dev, rects, offsets)."""
        import torch
        boxes = np.ascontiguousarray(np.asarray(boxes, dtype=np.int64).reshape(-1, 6))
        flat = [np.asarray(v, dtype=np.float32).ravel(order="F") for v in values]
        voff = np.zeros(len(flat) + 1, dtype=np.int64)
        voff[1:] = np.cumsum([f.size for f in flat])
        vals = torch.from_numpy(np.concatenate(flat) if flat else np.zeros(1, np.float32)).cuda()
        rects, offs = self.rects(boxes)
        out = torch.empty(int(max(offs[-1], 1)), dtype=torch.float32, device="cuda")
        return boxes, vals, voff, out, rects, offs

    def project(self, boxes, vals, voff, out, use_graph: bool = True):
        """Run the batched projection into ``out`` (device tensor)."""
        import torch
        P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        torch.cuda.synchronize()
        check(lib().sid_library_forward_device(self.psf._h, boxes.shape[0], P(boxes), ctypes.c_void_p(vals.data_ptr()),
                                               P(voff), i64(*self.sensor), ctypes.c_void_p(out.data_ptr()),
                                               1 if use_graph else 0,
                                               ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        return out

    def stats(self):
        buf = (ctypes.c_int64 * 4)()
        check(lib().sid_library_stats(self.psf._h, buf))
        return dict(zip(["rows", "items", "macs", "acc_elems"], list(buf)))

    @staticmethod
    def patch(out, rects, offs, s):
        """Seed s's patch as an (sx1-sx0+1, sy1-sy0+1) array (MATLAB indexing)."""
        r = rects[s]
        w, h = int(r[1] - r[0] + 1), int(r[3] - r[2] + 1)
        flat = out[int(offs[s]):int(offs[s + 1])]
        if hasattr(flat, "cpu"):
            flat = flat.cpu().numpy()
        return flat.reshape((w, h), order="F")


def kernel_boxes(boxes, sensor, kdims):
    """Boxes grown by a kernel's reach (sid_library_kernel_boxes): (n x 6 boxes,
    n + 1 value offsets)."""
    boxes = np.ascontiguousarray(np.asarray(boxes, dtype=np.int64).reshape(-1, 6))
    ns = boxes.shape[0]
    out = np.zeros_like(boxes)
    voff = np.zeros(ns + 1, dtype=np.int64)
    P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    check(lib().sid_library_kernel_boxes(ns, P(boxes), i64(*sensor), i64(*kdims), P(out), P(voff)))
    return out, voff


def convn_seeds(boxes, values, sensor, kernel):
    """convn(Volume, kernel, 'same') of each seed's volume (zero outside its
    box, src/generate_LFM_library_GPU.m:70-85,98) on the grown boxes
    (sid_library_convn_seeds, FP32): returns (grown boxes, per-seed values as
    column-major arrays)."""
    k = np.asfortranarray(np.asarray(kernel, dtype=np.float32))
    if k.ndim == 2:
        k = k[:, :, None]
    kd = tuple(int(v) for v in k.shape)
    boxes = np.ascontiguousarray(np.asarray(boxes, dtype=np.int64).reshape(-1, 6))
    flat = [np.asarray(v, dtype=np.float32).ravel(order="F") for v in values]
    voff = np.zeros(len(flat) + 1, dtype=np.int64)
    voff[1:] = np.cumsum([f.size for f in flat])
    vin = np.ascontiguousarray(np.concatenate(flat) if flat else np.zeros(1, np.float32))
    grown, goff = kernel_boxes(boxes, sensor, kd)
    vout = np.zeros(max(int(goff[-1]), 1), dtype=np.float32)
    P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    check(lib().sid_library_convn_seeds(boxes.shape[0], P(boxes), vin.ctypes.data_as(ctypes.c_void_p), P(voff),
                                        i64(*sensor), k.ctypes.data_as(ctypes.c_void_p), i64(*kd),
                                        vout.ctypes.data_as(ctypes.c_void_p)))
    out = []
    for s in range(boxes.shape[0]):
        g = grown[s]
        shape = (int(g[1] - g[0] + 1), int(g[3] - g[2] + 1), int(g[5] - g[4] + 1))
        out.append(vout[int(goff[s]):int(goff[s + 1])].reshape(shape, order="F"))
    return grown, out


def generate_LFM_library_GPU(recon, centers, neur_id, psf, options):
    """src/generate_LFM_library_GPU.m:1-113: row s of the single
    N x prod(image_size) result is the sensor image of centre s's sub-volume —
    the mean of the reconstructions listed in neur_id(s, :) on its clipped box
    (:68-85), convolved with generate_kernel(ker_shape, ker_param) when
    options has ker_param (:54-56,95-101) — projected with planes z0..z1 of the
    handle's PSF (:87-95), all seeds in one batched launch.  recon: list of
    [X, Y, Nz] arrays; centers: N x 3 (1-based); options: dict with image_size,
    axial, neur_rad and optionally ker_shape (default 'ball') and ker_param."""
    from .recon import generate_kernel
    X, Y = (int(v) for v in np.asarray(options["image_size"]).ravel()[:2])
    r = int(options["neur_rad"])
    dz = int(math.ceil(r / float(options["axial"])))
    nz = psf.Nz
    centers = np.asarray(centers, dtype=np.float64).reshape(-1, 3)
    neur_id = np.asarray(neur_id).reshape(centers.shape[0], -1)
    boxes, values = [], []
    for s in range(centers.shape[0]):
        b = placement(centers[s], r, dz, (X, Y), nz)
        ids = np.flatnonzero(neur_id[s])
        acc = np.zeros((b[1] - b[0] + 1, b[3] - b[2] + 1, b[5] - b[4] + 1))
        for k in ids:
            acc += np.asarray(recon[k])[b[0] - 1:b[1], b[2] - 1:b[3], b[4] - 1:b[5]]
        boxes.append(b)
        values.append((acc / len(ids)).astype(np.float32) if len(ids) else acc.astype(np.float32))
    if "ker_param" in options:
        kern = generate_kernel(options.get("ker_shape", "ball"), options["ker_param"])
        boxes, values = convn_seeds(boxes, values, (X, Y), kern)
    boxes = np.ascontiguousarray(np.asarray(boxes, dtype=np.int64).reshape(-1, 6))
    lb = Library(psf, (X, Y))
    rects, offs = lb.rects(boxes)
    flat = [np.asarray(v, dtype=np.float32).ravel(order="F") for v in values]
    voff = np.zeros(len(flat) + 1, dtype=np.int64)
    voff[1:] = np.cumsum([f.size for f in flat])
    vin = np.ascontiguousarray(np.concatenate(flat) if flat else np.zeros(1, np.float32))
    patches = np.zeros(max(int(offs[-1]), 1), dtype=np.float32)
    P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    check(lib().sid_library_forward(psf._h, boxes.shape[0], P(boxes), vin.ctypes.data_as(ctypes.c_void_p), P(voff),
                                    i64(X, Y), patches.ctypes.data_as(ctypes.c_void_p)))
    fm = np.zeros((boxes.shape[0], X * Y), dtype=np.float32)
    for s in range(boxes.shape[0]):
        rc = rects[s]
        img = np.zeros((X, Y), dtype=np.float32)
        img[rc[0] - 1:rc[1], rc[2] - 1:rc[3]] = Library.patch(patches, rects, offs, s)
        fm[s] = img.ravel(order="F")
    return fm


def generate_LFM_library_CPU(centers, psf, r, dim, m, kernel=None):
    """src/generate_LFM_library_CPU.m:1 on the device: the L2-normalised
    footprint of every centre (a ball of radius ``r`` with axial radius
    ``r / dim[2]``, or the nonzeros of ``kernel``) in a volume of lateral size
    ``m[:2]``, projected with ``psf`` (made without CAindex).

    Returns (row_ptr, cols, vals): the rows of the reference's
    ``sparse(I, J, S, N, prod(m(1:2)))`` in compressed-row form (0-based
    column-major pixel indices, ascending per row; FP64 values).
    ``scipy.sparse.csr_matrix((vals, cols, row_ptr), shape=(N, m[0]*m[1]))``
    builds the matrix."""
    c = np.asfortranarray(np.asarray(centers, dtype=np.float64).reshape(-1, 3))
    ns = c.shape[0]
    dm = (ctypes.c_double * 3)(*[float(v) for v in (dim if dim is not None else (1.0, 1.0, 1.0))])
    mm = i64(int(m[0]), int(m[1]))
    if kernel is not None:
        k = np.asfortranarray(np.asarray(kernel, dtype=np.float32))
        if k.ndim == 2:
            k = k[:, :, None]
        kp, kd = k.ctypes.data_as(ctypes.c_void_p), i64(*k.shape)
    else:
        kp, kd = None, None
    rv = float(r) if r is not None else 0.0
    bound = ctypes.c_int64(0)
    cp = c.ctypes.data_as(ctypes.c_void_p)
    check(lib().sid_library_nsf_bound(psf._h, ns, cp, rv, dm, mm, kp, kd, ctypes.byref(bound)))
    row_ptr = np.zeros(ns + 1, dtype=np.int64)
    cols = np.zeros(max(bound.value, 1), dtype=np.int64)
    vals = np.zeros(max(bound.value, 1), dtype=np.float64)
    P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    check(lib().sid_library_nsf(psf._h, ns, cp, rv, dm, mm, kp, kd, P(row_ptr), P(cols),
                                vals.ctypes.data_as(ctypes.c_void_p), int(bound.value)))
    nnz = int(row_ptr[-1])
    return row_ptr, cols[:nnz].copy(), vals[:nnz].copy()


def workload_seeds(wl):
    """Boxes and Gaussian-blob values of the batched-library workload: seeds
    with lateral box side in [box_lo, box_hi] (radius side // 2) and depth in
    [dz_lo, dz_hi] planes (half-depth depth // 2), placed by ``placement``."""
    from .synth import make_seeds
    sd = make_seeds(wl)
    boxes, vals = [], []
    for k in range(len(sd["box"])):
        r = int(sd["box"][k]) // 2
        dz = int(sd["depth"][k]) // 2
        c = sd["center"][k]
        b = placement(c, r, dz, (wl.X, wl.Y), wl.nz)
        boxes.append(b)
        vals.append(gaussian_blob(b, c, sd["sigma"]))
    return np.asarray(boxes, dtype=np.int64), vals
