#!/usr/bin/env python3
"""Benchmark: ISRA/RL reconstruction throughput of the light-field projector.

Workload (default ``rl_full``): 1005 x 1005 px sensor, Nnum = 15 (67 x 67
lenslets), 51 planes at z = -25..25, 345 px PSF array with per-plane
footprint up to 345 px, 3000 seeded ball neurons, Poisson image at a
1000-photon peak, reconstruct_S pre-processing, 8 ISRA iterations
(deconvRL basic, lambda = 0) entirely on one MI355X.

A "step" is one 8-iteration reconstruction (sid_deconv_rl_device); ``value``
is RL iterations per second of the BF16 hi/lo split MFMA path, with the
plain FP32 path timed in the same run.  Inputs are resident in HBM before
the timed region.  Errors are rel-L2 against the FP64 device path after
each of the 8 iterations.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

PEAK = {"fp32": 157.3e12, "bf16x2": 2.5e15, "bf16x3": 2.5e15, "fp64": 78.6e12}
HBM_PEAK = 8.0e12
PEAK_FP16 = 2.5e15    # dense FP16 MFMA (the tiled FFT path's spectra GEMMs)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(wl, H_host, htf_host):
    """Reference CPU path restated in C (oracle/liboracle.so: forwardProjectACC
    in single precision, backwardProjectACC_new stored double, CAindex crop,
    deconvRL's basic update), one whole RL iteration of the full workload on
    all host cores, started from the device's Htf (the input of the timed GPU
    iterations); plus the small workload end to end."""
    from oracle.build_oracle import build as build_oracle
    from sid_amd import synth
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle", "liboracle.so")
    if not os.path.exists(path):
        build_oracle()
    L = ctypes.CDLL(path)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    X, Y, nz = wl.X, wl.Y, wl.nz
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    # H_host holds MATLAB's column-major [M, M, n, n, Nz] memory; htf_host [X, Y, Nz]
    hd = (ctypes.c_int64 * 5)(wl.M, wl.M, wl.n, wl.n, nz)
    ca = np.asfortranarray(synth.ca_index(wl.zs, wl.M).astype(np.int32))
    htf = np.ascontiguousarray(htf_host, dtype=np.float64)
    Xg = htf.copy()
    img = np.zeros((X, Y), dtype=np.float32, order="F")
    D = np.zeros_like(htf)
    t0 = time.perf_counter()
    L.orc_isra_f32(P(H_host), hd, P(ca), P(htf), P(Xg), ctypes.c_int64(X), ctypes.c_int64(Y), ctypes.c_int64(nz),
                   ctypes.c_int(1), P(img), P(D), ctypes.c_int(threads))
    t1 = time.perf_counter()
    out = {"value": 1.0 / (t1 - t0), "unit": "RL iter/s", "cores": threads, "kind": "port", "cpu_model": cpu_model(),
           "sample": (f"one whole RL iteration of the full workload ({X}x{Y} px, {nz} planes, CAindex crops): "
                      f"forwardProjectACC (single) + backwardProjectACC_new (double) + the deconvRL update, "
                      f"restated in C (OpenMP, -O3), {t1 - t0:.2f} s from the device's Htf")}
    out["rl_small"] = cpu_rl_small(L, threads)
    return out


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_rl_small(L, threads):
    """The small workload in full on the host: 255 x 255 px, 11 planes, 105 px
    kernels, reconstruct_S pre-processing, 8 ISRA iterations of the reference
    CPU path's arithmetic (single-precision forward, double-stored backward,
    CAindex crops), timed whole."""
    from sid_amd import recon, synth
    wl = synth.WORKLOADS["rl_small"]
    H, CA = synth.make_psf(wl.n, wl.M, wl.zs, dtype=np.float32)
    V = synth.make_volume(wl.X, wl.Y, wl.nz, wl.n_balls, wl.r_lo, wl.r_hi, wl.seed, wl.background)
    X, Y, nz = wl.X, wl.Y, wl.nz
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    hd = (ctypes.c_int64 * 5)(*H.shape)
    ca = np.asfortranarray(CA.astype(np.int32))
    O = np.zeros((X, Y), dtype=np.float32, order="F")
    L.orc_forward_f32(P(H), hd, P(ca), P(np.asfortranarray(V.astype(np.float64))), ctypes.c_int64(X),
                      ctypes.c_int64(Y), ctypes.c_int64(nz), P(O), ctypes.c_int(threads))
    img = synth.poisson_image(O, wl.peak_photons, wl.seed)
    y, _ = recon.preprocess(img)
    yf = np.asfortranarray(y.astype(np.float32))
    t0 = time.perf_counter()
    htf = np.zeros((X, Y, nz), dtype=np.float64, order="F")
    L.orc_backward_f32(P(H), hd, P(ca), P(yf), ctypes.c_int64(X), ctypes.c_int64(Y), ctypes.c_int64(nz), P(htf),
                       ctypes.c_int(threads))
    Xg = htf.copy(order="F")
    img_w = np.zeros((X, Y), dtype=np.float32, order="F")
    D = np.zeros((X, Y, nz), dtype=np.float64, order="F")
    t1 = time.perf_counter()
    L.orc_isra_f32(P(H), hd, P(ca), P(htf), P(Xg), ctypes.c_int64(X), ctypes.c_int64(Y), ctypes.c_int64(nz),
                   ctypes.c_int(wl.iters), P(img_w), P(D), ctypes.c_int(threads))
    t2 = time.perf_counter()
    return {"workload": f"rl_small: {X}x{Y} px, {nz} planes, M={wl.M}, {wl.iters} ISRA iterations",
            "iter_per_s": round(wl.iters / (t2 - t1), 4), "s_isra": round(t2 - t1, 3),
            "s_htf_backward": round(t1 - t0, 3), "proj_per_s": round(2 * wl.iters / (t2 - t1), 3)}


def projection_rates(dev, names=("rl_small",), precisions=("fp32", "bf16x2"), reps=20):
    """Forward and backward projections per second on device buffers (the
    autotuned plan of each workload's handle)."""
    import torch
    import sid_amd
    from sid_amd import synth
    out = {}
    for name in names:
        wl = synth.WORKLOADS[name]
        H, _ = synth.make_psf(wl.n, wl.M, wl.zs, xp=torch, device=dev)
        psf = sid_amd.PSF(H)
        del H
        V = torch.rand((wl.nz, wl.Y, wl.X), device=dev)
        O = torch.rand((wl.Y, wl.X), device=dev)
        res = {}
        for prec in precisions:
            psf.set_precision(prec)
            psf.forward(V), psf.backward(O)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                psf.forward(V)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for _ in range(reps):
                psf.backward(O)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            res[prec] = {"forward_per_s": round(reps / (t1 - t0), 1), "backward_per_s": round(reps / (t2 - t1), 1)}
        out[name] = res
        log(f"projections {name}: " + ", ".join(f"{k} fwd {v['forward_per_s']:.0f}/s bwd {v['backward_per_s']:.0f}/s"
                                                   for k, v in res.items()))
        psf.close()
        del V, O
        torch.cuda.empty_cache()
    return out


def grad_bench(dev, reps=3):
    """Differentiation row: Poisson NLL of a batch of 16 frames (495 x 495 px,
    31 planes z=-15..15, 225 px kernels, 40 balls each) and its gradients with
    respect to the 16 volumes and the PSF, FP32 direct path.  Algorithmic work
    per frame: forward, backward and PSF-weight-gradient GEMMs of
    2 * X * Y * sum_c M_c^2 flop each."""
    import torch
    import sid_amd
    from sid_amd import grad, synth
    wl = synth.WORKLOADS["grad"]
    H, _ = synth.make_psf(wl.n, wl.M, wl.zs, xp=torch, device=dev)
    psf = sid_amd.PSF(H, method="direct")
    del H
    psf.set_precision("fp32")
    msum = sum((hi - lo + 1) ** 2 for lo, hi in psf.info()["support"] if hi >= lo)
    nb = wl.batch
    vols, imgs = [], []
    for f in range(nb):
        V = synth.make_volume(wl.X, wl.Y, wl.nz, wl.n_balls, wl.r_lo, wl.r_hi, wl.seed + f, wl.background)
        vols.append(torch.from_numpy(np.ascontiguousarray(V.transpose(2, 1, 0))).to(dev))
        img = synth.poisson_image(psf.forward(vols[-1]).cpu().numpy().T, wl.peak_photons, wl.seed + f)
        imgs.append(torch.from_numpy(np.ascontiguousarray(img.T.astype(np.float32))).to(dev))
    V, y = torch.stack(vols), torch.stack(imgs)
    del vols, imgs
    times = {}
    for name, gv, gh in (("nll_gV", True, False), ("nll_gV_gH", True, True)):
        grad.poisson_nll_grad(psf, V, y, gv, gh)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            nll, _, _ = grad.poisson_nll_grad(psf, V, y, gv, gh)
        torch.cuda.synchronize()
        times[name] = (time.perf_counter() - t0) / reps
    fl_gemm = 2.0 * wl.X * wl.Y * msum
    t, tv = times["nll_gV_gH"], times["nll_gV"]
    # forward / backward projections per second at this workload (FP32 direct)
    psf.forward(V[0]), psf.backward(y[0])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for f in range(nb):
        psf.forward(V[f])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for f in range(nb):
        psf.backward(y[f])
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    proj = {"forward_per_s": round(nb / (t1 - t0), 1), "backward_per_s": round(nb / (t2 - t1), 1)}
    res = {"frames": nb, "X": wl.X, "Y": wl.Y, "Nz": wl.nz, "M": wl.M, "precision": "fp32",
           "ms_nll_gV_gH": round(t * 1e3, 2), "ms_nll_gV": round(tv * 1e3, 2),
           "batches_per_s": round(1.0 / t, 3), "frames_per_s": round(nb / t, 2),
           "tflops": round(3 * nb * fl_gemm / t / 1e12, 2),
           "frac_of_fp32_peak": round(3 * nb * fl_gemm / t / PEAK["fp32"], 4),
           "psf_grad_tflops": round(nb * fl_gemm / max(t - tv, 1e-9) / 1e12, 2), "nll": nll,
           "projections_fp32_direct": proj}
    log(f"gradients ({nb} frames {wl.X}x{wl.Y}x{wl.nz}, M={wl.M}): NLL+dV+dH {t * 1e3:.1f} ms "
        f"({res['tflops']:.1f} TFLOP/s), NLL+dV {tv * 1e3:.1f} ms")
    del psf, V, y
    torch.cuda.empty_cache()
    return res


def convn_transform_bytes(X, Y, Z, kshape):
    """HBM bytes of one transform-form convn (the five passes of sid_convn_fft.hip:
    x forward, y forward, z sum, y inverse, x inverse) at its transform lengths:
    the smallest of 128/256/512/768/1024/2048 covering X + kx//2 and Y + ky//2."""
    kx, ky, kz = kshape
    ln = [L for L in (128, 256, 512, 768, 1024, 2048)]
    Nx = next(L for L in ln if L >= max(X + kx // 2, kx))
    Ny = next(L for L in ln if L >= max(Y + ky // 2, ky))
    Hx, Yp = Nx // 2 + 1, -(-Y // 16) * 16
    vol, col = 4 * X * Y * Z, 8 * Z * Hx
    return (vol + col * Yp) + (col * Y + col * Ny) + (2 * col * Ny + 8 * kz * Hx * Ny) + (col * Ny + col * Yp) \
        + (col * Yp + vol)


def library_bench(psf, wl_lib=None, steps=3):
    """Batched library workload: 2500 seeds (lateral box 7-21 px, depth 3-9
    planes, Gaussian blob sigma 2 px) in the 1005 x 1005 x 51 geometry, one
    persistent hipGraph-captured launch per batch; a few seeds are checked
    against the full-sensor forward projection of their sub-volume."""
    import torch
    from sid_amd import library, synth
    wl = synth.WORKLOADS["library"]
    psf.set_precision("fp32")
    lib = library.Library(psf, (wl.X, wl.Y))
    t = time.perf_counter()
    boxes, vals = library.workload_seeds(wl)
    b, v, voff, out, rects, offs = lib.prepare(boxes, vals)
    lib.project(b, v, voff, out, use_graph=True)          # plan + capture + first launch
    torch.cuda.synchronize()
    log(f"library: {len(boxes)} seeds planned and captured in {time.perf_counter() - t:.1f} s")
    t0 = time.perf_counter()
    for _ in range(steps):
        lib.project(b, v, voff, out, use_graph=True)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    st = lib.stats()
    errs = []
    for s in (0, len(boxes) // 2, len(boxes) - 1):
        x0, x1, y0, y1, z0, z1 = (int(q) for q in boxes[s])
        V = torch.zeros((wl.nz, wl.Y, wl.X), dtype=torch.float32, device="cuda")
        V[z0 - 1:z1, y0 - 1:y1, x0 - 1:x1] = torch.from_numpy(np.ascontiguousarray(vals[s].transpose(2, 1, 0))).cuda()
        full = psf.forward(V).cpu().numpy().T.astype(np.float64)
        r = rects[s]
        ref = full[r[0] - 1:r[1], r[2] - 1:r[3]]
        got = library.Library.patch(out, rects, offs, s).astype(np.float64)
        errs.append(float(np.linalg.norm(got - ref) / np.linalg.norm(ref)))
    res = {"seeds": int(len(boxes)), "ms_per_batch": round(dt * 1e3, 3), "seeds_per_s": round(len(boxes) / dt, 1),
           "mfma_tflops": round(2.0 * st["macs"] / dt / 1e12, 2), "work_items": st["items"],
           "patch_pixels": int(offs[-1]), "spot_rel_l2_vs_full_sensor_fp32": errs}
    log(f"library: {res['seeds_per_s']:.0f} seeds/s ({res['ms_per_batch']:.2f} ms per batch of {len(boxes)}), "
        f"spot-check rel-L2 {max(errs):.2e}")
    del out, v
    return res


def rows_bench(psf, wl, htf, prec, iters, dev):
    """The solver / wrapper / producer rows around the projector, each timed at
    the full workload in the headline precision:
    LS_deconv (reconstruct_S 'fast_nnls'), the neuron-kernel wrapping (a
    find_kernel-sized 17 x 17 x 5 ball) and the movie std image (1005^2 px x
    1000 frames, bg_iter 2, max-normalised)."""
    import torch
    import sid_amd
    from sid_amd import movie, recon
    out = {}
    psf.set_precision(prec)
    lam = min(recon.quantile_type5(htf.cpu().numpy(), 1 - 1e-4), 0.1)
    opts = {"max_iter": iters, "lamb_L1": lam}
    psf.ls_deconv(htf, opts)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    psf.ls_deconv(htf, opts)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out["ls_deconv"] = {"iter_per_s": round(iters / dt, 3), "ms_per_solve": round(dt * 1e3, 2), "max_iter": iters,
                        "projections_per_iter": 3, "lamb_L1": lam}
    log(f"LS_deconv ({prec}): {iters / dt:.2f} iter/s ({dt * 1e3:.0f} ms for {iters} iterations)")
    k = recon.generate_kernel("ball", (8, 2))
    kt = torch.from_numpy(np.ascontiguousarray(k.T)).to(dev)
    vol = torch.rand((wl.nz, wl.Y, wl.X), device=dev)
    cprec = prec if prec in ("bf16x2", "bf16x3") else "fp32"
    sid_amd.convn_same(vol, kt, precision=cprec)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        sid_amd.convn_same(vol, kt, precision=cprec)
    torch.cuda.synchronize()
    tc = (time.perf_counter() - t0) / reps
    sid_amd.convn_same(vol, kt, precision="fp32")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        sid_amd.convn_same(vol, kt, precision="fp32")
    torch.cuda.synchronize()
    tc32 = (time.perf_counter() - t0) / reps
    # the transform form against the direct forms (SID_CONVN_FFT=0): the BF16 MFMA
    # form and the FP32 direct sum
    tcm = None
    os.environ["SID_CONVN_FFT"] = "0"
    try:
        if cprec != "fp32":
            sid_amd.convn_same(vol, kt, precision=cprec)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                sid_amd.convn_same(vol, kt, precision=cprec)
            torch.cuda.synchronize()
            tcm = (time.perf_counter() - t0) / reps
        sid_amd.convn_same(vol, kt, precision="fp32")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            sid_amd.convn_same(vol, kt, precision="fp32")
        torch.cuda.synchronize()
        tcd = (time.perf_counter() - t0) / reps
    finally:
        del os.environ["SID_CONVN_FFT"]
    taps = int(np.prod(k.shape))
    fl = 2.0 * wl.X * wl.Y * wl.nz * taps
    psf.set_kernel(k)
    psf.deconv_rl(htf, 1, htf)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    psf.deconv_rl(htf, iters, htf)
    torch.cuda.synchronize()
    dtw = time.perf_counter() - t0
    psf.set_kernel(None)
    form = "transform" if wl.X + k.shape[0] // 2 <= 2048 and wl.Y + k.shape[1] // 2 <= 2048 and k.shape[2] <= 8 \
        else ("mfma" if cprec != "fp32" else "direct")
    hbm = convn_transform_bytes(wl.X, wl.Y, wl.nz, k.shape) if form == "transform" else None
    out["neuron_kernel"] = {"kernel": list(k.shape), "precision": cprec, "convn_form": form,
                            "convn_hbm_gb": round(hbm / 1e9, 3) if hbm else None,
                            "convn_hbm_tbps": round(hbm / tc / 1e12, 3) if hbm else None,
                            "convn_frac_of_hbm": round(hbm / tc / 8e12, 4) if hbm else None,
                            "convn_mfma_ms": round(tcm * 1e3, 3) if tcm else None,
                            "convn_ms": round(tc * 1e3, 3),
                            "convn_tflops": round(fl / tc / 1e12, 2), "convn_fp32_ms": round(tc32 * 1e3, 3),
                            "convn_fp32_direct_ms": round(tcd * 1e3, 3),
                            "convn_fp32_direct_tflops": round(fl / tcd / 1e12, 2), "fp32_vector_peak_tflops": 78.6,
                            "convn_frac_of_peak": round(fl / tc / PEAK[cprec], 4),
                            "convn_fp32_direct_frac_of_vector_peak": round(fl / tcd / 78.6e12, 4),
                            "rl_iter_per_s_wrapped": round(iters / dtw, 3)}
    rate = f"{hbm / tc / 1e12:.2f} TB/s over its {hbm / 1e9:.2f} GB of passes, " if hbm else ""
    log(f"neuron kernel {k.shape}: convn {cprec} {form} {tc * 1e3:.2f} ms ({fl / tc / 1e12:.1f} TFLOP/s algorithmic, "
        f"{rate}MFMA form {tcm * 1e3 if tcm else float('nan'):.2f} ms), "
        f"FP32 {tc32 * 1e3:.2f} ms (direct sum {tcd * 1e3:.2f} ms); wrapped RL {iters / dtw:.2f} iter/s")
    del vol
    P, T = wl.X * wl.Y, 1000
    Y = torch.rand((T, P), device=dev)
    movie.background_and_std(Y, bg_iter=2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    movie.background_and_std(Y, bg_iter=2)
    torch.cuda.synchronize()
    tm = time.perf_counter() - t0
    passes = 4   # 2 column passes (the first also finds the maximum), 2 row passes
    out["std_image"] = {"pixels": P, "frames": T, "bg_iter": 2, "ms": round(tm * 1e3, 2),
                        "stream_gbps": round(passes * P * T * 4 / tm / 1e9, 1), "passes": passes,
                        "frac_of_hbm": round(passes * P * T * 4 / tm / HBM_PEAK, 4)}
    log(f"std image {P} px x {T} frames: {tm * 1e3:.1f} ms ({passes * P * T * 4 / tm / 1e9:.0f} GB/s over {passes} passes)")
    del Y
    torch.cuda.empty_cache()
    return out


def pmc_summary(fetch_dir, write_dir, bench_out, dest):
    """Summarise two rocprofv3 counter runs of the same bench command (FETCH_SIZE, then WRITE_SIZE, each with
    --profile-copy) into the traffic summary the roofline reads: per kernel name the scaled bytes of its
    last dispatches, and the tiled path's spectra GEMMs of the last forward + backward projection pair.
    The scales come from the 1 GiB device copy (1 GiB read + 1 GiB written) in the same runs."""
    import csv
    import glob

    def load(d, counter):
        rows = []
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            rows += [r for r in csv.DictReader(open(f)) if r["Counter_Name"] == counter]
        disp = {}
        for r in rows:
            k = int(r["Dispatch_Id"])
            e = disp.setdefault(k, {"name": r["Kernel_Name"], "v": 0.0,
                                    "ms": (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6})
            e["v"] += float(r["Counter_Value"]) * 1024.0   # KiB per dispatch
        return [disp[k] for k in sorted(disp)]

    fetch, write = load(fetch_dir, "FETCH_SIZE"), load(write_dir, "WRITE_SIZE")
    copy_f = [d for d in fetch if "copyBuffer" in d["name"]]
    copy_w = [d for d in write if "copyBuffer" in d["name"]]
    if not copy_f or not copy_w:
        raise SystemExit("no 1 GiB copy dispatch in the counter runs (run the bench with --profile-copy)")
    fs, ws = float(1 << 30) / copy_f[0]["v"], float(1 << 30) / copy_w[0]["v"]
    line = [l for l in open(bench_out) if l.startswith("{")][-1]
    res = json.loads(line)
    head = res["config"]["method"]
    hr = res["methods"][head]
    kern = {}
    for lst, key, sc in ((fetch, "read_bytes", fs), (write, "write_bytes", ws)):
        names = {}
        for d in lst:
            names.setdefault(d["name"], []).append(d)
        for nm, ds in names.items():
            last = ds[-8:]
            e = kern.setdefault(nm, {})
            e[key] = sum(d["v"] for d in last) / len(last) * sc
            e["ms"] = sum(d["ms"] for d in last) / len(last)
    out = {"workload": res["config"]["workload"].split(":")[0], "precision": head, "method": res["config"]["projector"],
           "spectra_gb": hr["spectra_gb"], "fetch_scale": fs, "write_scale": ws, "kernels": kern}
    if hr.get("fft_tiled"):
        ng = len(hr["fft_groups"])
        gf = [d for d in fetch if "k_tf_gemm" in d["name"]][-2 * ng:]
        gw = [d for d in write if "k_tf_gemm" in d["name"]][-2 * ng:]
        out["fft_groups"] = hr["fft_groups"]
        out["tf_gemm_read_bytes"] = [sum(d["v"] for d in gf[:ng]) * fs, sum(d["v"] for d in gf[ng:]) * fs]
        out["tf_gemm_write_bytes"] = [sum(d["v"] for d in gw[:ng]) * ws, sum(d["v"] for d in gw[ng:]) * ws]
        out["tf_gemm_bytes"] = sum(out["tf_gemm_read_bytes"]) + sum(out["tf_gemm_write_bytes"])
        out["tf_gemm_algorithmic_bytes"] = [g * 1e9 for g in hr["tf_gemm_gb"]]
        out["tf_gemm_ms"] = [sum(d["ms"] for d in gf[:ng]), sum(d["ms"] for d in gf[ng:])]
    json.dump(out, open(dest, "w"), indent=1)
    print(json.dumps({k: v for k, v in out.items() if k != "kernels"}, indent=1))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--iters", type=int, default=8, help="RL iterations per step")
    ap.add_argument("--workload", default="rl_full")
    ap.add_argument("--precisions", default="fp32,bf16x2", help="first is the baseline, last is the headline")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-ref", action="store_true", help="skip the FP64 error reference")
    ap.add_argument("--no-library", action="store_true", help="skip the batched-library measurement")
    ap.add_argument("--no-rows", action="store_true",
                    help="skip the LS_deconv / neuron-kernel / std-image measurements")
    ap.add_argument("--profile-copy", action="store_true",
                    help="run one 1 GiB device copy first (the reference kernel for scaling PMC byte counters)")
    ap.add_argument("--method", default="auto", choices=["auto", "direct", "fft"],
                    help="projector method per plane (auto = autotuned direct-vs-FFT choice)")
    ap.add_argument("--grad-only", action="store_true",
                    help="run only the differentiation row (NLL + volume and PSF gradients, 16 frames 495^2 x 31)")
    ap.add_argument("--no-host", action="store_true",
                    help="skip the host-buffer leg (sid_deconv_rl with PCIe transfers, one 8-iteration call)")
    ap.add_argument("--no-direct", action="store_true",
                    help="skip the direct-only leg (every plane on the MFMA GEMM, one step per precision)")
    ap.add_argument("--clean-exit", action="store_true",
                    help="end through the interpreter's normal teardown (profilers such as rocprofv3 write their "
                         "results there); by default the process leaves right after the result line")
    ap.add_argument("--fft-planes", type=int, default=-1,
                    help="pin the K widest planes to the FFT path and the rest to the direct GEMM (reproduces an "
                         "autotuned split in profiler runs, whose serialised kernels would mislead the autotuner)")
    ap.add_argument("--pmc-summary", nargs=3, metavar=("FETCH_DIR", "WRITE_DIR", "BENCH_OUT"),
                    help="summarise two rocprofv3 counter runs (FETCH_SIZE, WRITE_SIZE) of one bench command and "
                         "its output line into profiles/traffic_rl_full.json, then exit (no GPU use)")
    args = ap.parse_args()
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary, os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                                                    "traffic_rl_full.json"))
        return args
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {args.gpus}): the projector is measured on one "
              f"MI355X; multi-GPU sharding of this path is a design note, not a measured run.", file=sys.stderr)
        sys.exit(2)

    import torch
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import sid_amd
    from sid_amd import recon, synth

    if sid_amd.device_count() < 1:
        print("no gfx950 device visible", file=sys.stderr)
        sys.exit(1)
    wl = synth.WORKLOADS[args.workload]
    dev = torch.device("cuda:0")
    if args.grad_only:
        print(json.dumps({"gradient": grad_bench(dev)}))
        return args
    if args.profile_copy:
        a = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
        log("reference copy: 1 GiB read + 1 GiB written")
    log(f"workload {wl.name}: X={wl.X} Y={wl.Y} n={wl.n} Nz={wl.nz} M={wl.M}")
    t = time.perf_counter()
    H, CA = synth.make_psf(wl.n, wl.M, wl.zs, xp=torch, device=dev)
    torch.cuda.synchronize()
    log(f"PSF generated on device in {time.perf_counter() - t:.1f} s ({H.numel() * 4 / 1e9:.2f} GB)")
    t = time.perf_counter()
    psf = sid_amd.PSF(H, method=args.method)
    del H
    torch.cuda.empty_cache()
    info = psf.info()
    if args.fft_planes >= 0:
        widths = [hi - lo + 1 for lo, hi in info["support"]]
        order = sorted(range(len(widths)), key=lambda c: -widths[c])
        pins = ["direct"] * len(widths)
        for c in order[:args.fft_planes]:
            pins[c] = "fft"
        psf.set_plane_methods(pins)
        log(f"pinned {args.fft_planes} widest planes to the FFT path")
    log(f"PSF handle built in {time.perf_counter() - t:.1f} s; taps_max={info['taps_max']} "
        f"weights {info['weight_bytes'] / 1e9:.2f} GB")
    V = synth.make_volume(wl.X, wl.Y, wl.nz, wl.n_balls, wl.r_lo, wl.r_hi, wl.seed, wl.background)
    Vt = torch.from_numpy(np.ascontiguousarray(V.transpose(2, 1, 0))).to(dev)
    clean = psf.forward(Vt)
    img = synth.poisson_image(clean.cpu().numpy().T, wl.peak_photons, wl.seed)
    y, nrm = recon.preprocess(img)
    yt = torch.from_numpy(np.ascontiguousarray(y.T.astype(np.float32))).to(dev)
    htf = psf.backward(yt)
    torch.cuda.synchronize()
    log("inputs ready (volume, Poisson image, Htf = B(y/nrm))")

    # algorithmic work per projection: X*Y*sum_c M_c^2 MACs (M_c = plane support side)
    msum = sum((hi - lo + 1) ** 2 for lo, hi in info["support"] if hi >= lo)
    flops_proj = 2.0 * wl.X * wl.Y * msum

    results = {}
    htfs = {}
    for prec in args.precisions.split(","):
        psf.set_precision(prec)
        # Xguess = Htf = B(img/nrm) in the same arithmetic as the iterations (reconstruct_S.m:192-196)
        htfs[prec] = psf.backward(yt)
        for _ in range(args.warmup):
            psf.deconv_rl(htfs[prec], args.iters, htfs[prec])
        torch.cuda.synchronize()
        tim = np.zeros(12)
        t_steps = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ts = time.perf_counter()
            psf.deconv_rl(htfs[prec], args.iters, htfs[prec])   # returns after its last iteration's timing events
            tim += np.array(psf.timing())
            t_steps.append(time.perf_counter() - ts)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ms_step = (t1 - t0) * 1e3 / args.steps
        n_launch = args.steps * args.iters
        t_f, t_b = tim[0] / n_launch, tim[2] / n_launch          # direct GEMM per launch
        t_ff, t_fb = tim[5] / n_launch, tim[6] / n_launch        # FFT path per projection
        t_gf, t_gb = tim[7] / n_launch, tim[8] / n_launch        # spectra GEMVs inside it
        inf = psf.info()
        meth = inf["plane_method"]
        sup = inf["support"]
        flops_direct = 2.0 * wl.X * wl.Y * sum((hi - lo + 1) ** 2 for (lo, hi), m in zip(sup, meth) if m == "direct")
        res = dict(iter_per_s=args.steps * args.iters / (t1 - t0), ms_per_step=ms_step,
                   median_step_iter_per_s=args.iters / float(np.median(t_steps)),
                   fwd_gemm_ms=t_f, bwd_gemm_ms=t_b, fwd_fft_ms=t_ff, bwd_fft_ms=t_fb,
                   fwd_gemv_ms=t_gf, bwd_gemv_ms=t_gb,
                   fwd_reduce_ms=tim[1] / n_launch, bwd_reduce_ms=tim[3] / n_launch,
                   planes_direct=sum(m == "direct" for m in meth), planes_fft=sum(m == "fft" for m in meth),
                   spectra_gb=inf["spectra_bytes"] / 1e9, autotune_ms=inf["autotune_ms"],
                   proj_per_s=2.0 / ((t_f + t_b + t_ff + t_fb + (tim[1] + tim[3]) / n_launch) * 1e-3))
        if t_f + t_b > 0 and flops_direct > 0:
            res["gemm_achieved_tflops"] = flops_direct / (0.5 * (t_f + t_b) * 1e-3) / 1e12
            res["gemm_frac_of_peak"] = res["gemm_achieved_tflops"] * 1e12 / PEAK[prec]
        res["fft_tiled"] = inf["fft_tile"][0] > 0
        if t_gf + t_gb > 0 and res["fft_tiled"]:
            # tiled FFT path: the spectra GEMMs of all plane groups, per projection (one launch per
            # group); bytes they must move at least: FP16 spectra once, input spectra read, results written
            gb_f, gb_b = inf["fft_gemm_bytes"]
            fl_f, fl_b = inf["fft_gemm_flops"]
            res["fft_groups"] = [list(g) for g in inf["fft_groups"]]
            res["tf_gemm_gb"] = [gb_f / 1e9, gb_b / 1e9]
            res["tf_gemm_stream_tbps"] = (gb_f + gb_b) / ((t_gf + t_gb) * 1e-3) / 1e12
            res["tf_gemm_frac_of_hbm"] = res["tf_gemm_stream_tbps"] * 1e12 / HBM_PEAK
            res["tf_gemm_mfma_tflops"] = (fl_f + fl_b) / ((t_gf + t_gb) * 1e-3) / 1e12
            res["tf_gemm_mfma_frac"] = res["tf_gemm_mfma_tflops"] * 1e12 / PEAK_FP16
        elif t_gf + t_gb > 0:
            # FFT path: each GEMV streams the spectra once (algorithmic bytes per launch)
            res["gemv_stream_tbps"] = inf["spectra_bytes"] / (0.5 * (t_gf + t_gb) * 1e-3) / 1e12
            res["gemv_frac_of_hbm"] = res["gemv_stream_tbps"] * 1e12 / HBM_PEAK
        # MFMA work of the direct planes: algorithmic (X*Y*M_c^2 per plane) against executed (K rows run)
        st8 = psf.plan_stats()
        pm_f, pm_b = psf.plane_macs()
        res["gemm_macs_executed"] = 0.5 * (st8["macs_fwd"] + st8["macs_bwd"])
        res["gemm_macs_algorithmic"] = flops_direct / 2.0
        res["plane_exec_over_alg"] = [round(0.5 * (pm_f[c] + pm_b[c]) / (wl.X * wl.Y * (hi - lo + 1) ** 2), 3)
                                      if m == "direct" else None for c, ((lo, hi), m) in enumerate(zip(sup, meth))]
        results[prec] = res
        log(f"{prec}: {res['iter_per_s']:.2f} RL iter/s, {ms_step:.1f} ms/step; per projection: direct GEMM "
            f"fwd {t_f:.2f} / bwd {t_b:.2f} ms on {res['planes_direct']} planes, FFT path fwd {t_ff:.2f} / "
            f"bwd {t_fb:.2f} ms on {res['planes_fft']} planes ({res['spectra_gb']:.1f} GB spectra)")

    # The tiled FFT path runs its plane groups in two stream chains side by side, so each spectra
    # GEMM shares the GPU with the other chain's kernels and its event time is not its own rate.
    # One more step of the headline precision with one chain gives the GEMMs' alone-rate for the
    # roofline (the headline value stays the two-chain run's).
    head0 = args.precisions.split(",")[-1]
    if results.get(head0, {}).get("fft_tiled"):
        psf.set_precision(head0)
        psf.set_overlap(False)
        psf.deconv_rl(htfs[head0], args.iters, htfs[head0])
        torch.cuda.synchronize()
        psf.timing()
        tim1 = np.zeros(12)
        reps = max(1, min(args.steps, 3))
        t0 = time.perf_counter()
        for _ in range(reps):
            psf.deconv_rl(htfs[head0], args.iters, htfs[head0])
            tim1 += np.array(psf.timing())
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        n1 = reps * args.iters
        hr0 = results[head0]
        gb = sum(hr0["tf_gemm_gb"]) * 1e9
        hr0["one_chain"] = {"iter_per_s": n1 / (t1 - t0), "fwd_gemm_ms": tim1[7] / n1, "bwd_gemm_ms": tim1[8] / n1,
                            "tf_gemm_stream_tbps": gb / ((tim1[7] + tim1[8]) / n1 * 1e-3) / 1e12}
        psf.set_overlap(True)
        log(f"{head0} with one stream chain: {hr0['one_chain']['iter_per_s']:.2f} RL iter/s; spectra GEMMs "
            f"{hr0['one_chain']['fwd_gemm_ms']:.3f} / {hr0['one_chain']['bwd_gemm_ms']:.3f} ms per projection "
            f"= {hr0['one_chain']['tf_gemm_stream_tbps']:.2f} TB/s")

    # errors against the FP64 device path (FP64 MFMA, float I/O), per output and per iteration
    rel = {}
    ref_name = None
    if not args.no_ref:
        psf.set_precision("fp64")
        ref_name = "fp64"
        t = time.perf_counter()
        O64 = psf.forward(Vt).double()
        h64 = psf.backward(yt)
        _, hist64 = psf.deconv_rl(h64, args.iters, h64, history=True)
        torch.cuda.synchronize()
        log(f"FP64 reference done in {time.perf_counter() - t:.1f} s")
        h64d = h64.double()
        for prec in htfs:
            psf.set_precision(prec)
            Op = psf.forward(Vt).double()
            _, h = psf.deconv_rl(htfs[prec], args.iters, htfs[prec], history=True)
            per_it = []
            max_abs = []
            for i in range(args.iters):
                a, b = h[i].double(), hist64[i].double()
                per_it.append(float((a - b).norm() / b.norm()))
                max_abs.append(float((a - b).abs().max() / b.abs().max()))
            rel[prec] = {"forward": float((Op - O64).norm() / O64.norm()),
                         "backward": float((htfs[prec].double() - h64d).norm() / h64d.norm()),
                         "rl_per_iter": per_it, "rl_max_abs_rel_per_iter": max_abs, "final": per_it[-1]}
            log(f"{prec} rel-L2 vs FP64: forward {rel[prec]['forward']:.2e} backward {rel[prec]['backward']:.2e} "
                f"RL per iteration " + " ".join(f"{e:.2e}" for e in per_it))
            del h
        del hist64

    if not args.no_library:
        out_lib = library_bench(psf, wl_lib=None, steps=max(args.steps, 3))
    proj_rates = projection_rates(dev) if not args.no_rows else None
    out_rows = None
    if not args.no_rows:
        head_prec = args.precisions.split(",")[-1]
        out_rows = rows_bench(psf, wl, htfs[head_prec], head_prec, args.iters, dev)
        out_rows["gradient"] = grad_bench(dev)
    head = args.precisions.split(",")[-1]
    base = args.precisions.split(",")[0]
    host_rate = None
    if not args.no_host:
        # the host-buffer entry (sid_deconv_rl, what a MATLAB MEX caller uses): Htf and X cross PCIe
        # from and to pageable host memory around the same 8 device iterations
        # (the C entry on preallocated column-major arrays, as a MEX gateway calls it)
        import ctypes
        from sid_amd._lib import check, i64, lib
        psf.set_precision(head)
        h_np = np.asfortranarray(htfs[head].cpu().numpy().transpose(2, 1, 0))
        x_np = np.empty_like(h_np, order="F")
        ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        for rep in range(2):   # the first call also sets up the staging buffers
            np.copyto(x_np, h_np)
            t0 = time.perf_counter()
            check(lib().sid_deconv_rl(psf._h, ptr(h_np), ptr(x_np), i64(wl.X, wl.Y, wl.nz), args.iters, None))
            dt = time.perf_counter() - t0
        host_rate = round(args.iters / dt, 3)
        del h_np, x_np
        log(f"host-buffer entry ({head}): {host_rate:.2f} RL iter/s including the PCIe transfers")
    direct = None
    if not args.no_direct and args.method == "auto":
        # every plane on the implicit GEMM: the MFMA path alone, per precision (one timed step each)
        direct, direct_roof = {}, {}
        psf.set_plane_methods(None)
        psf.set_method("direct")
        for prec in args.precisions.split(","):
            psf.set_precision(prec)
            psf.deconv_rl(htfs[prec], args.iters, htfs[prec])
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            psf.deconv_rl(htfs[prec], args.iters, htfs[prec])
            torch.cuda.synchronize()
            direct[prec] = round(args.iters / (time.perf_counter() - t0), 3)
            # the GEMM alone over every plane: algorithmic and executed MFMA rate per launch
            tm = psf.timing()
            t_launch = 0.5 * (tm[0] + tm[2]) / args.iters * 1e-3
            st8 = psf.plan_stats()
            prods = {"bf16x2": 2, "bf16x3": 3}.get(prec, 1)
            ex = 2.0 * 0.5 * (st8["macs_fwd"] + st8["macs_bwd"]) * prods
            direct_roof[prec] = {"ms_per_launch": round(t_launch * 1e3, 3),
                                 "achieved_tflops": round(flops_proj / t_launch / 1e12, 2),
                                 "frac": round(flops_proj / t_launch / PEAK[prec], 4),
                                 "executed_mfma_frac": round(ex / t_launch / PEAK[prec], 4),
                                 "executed_over_algorithmic": round(ex / prods / flops_proj, 3)}
        psf.set_method("auto")
        log("direct-only (every plane on the MFMA GEMM): " + ", ".join(f"{k} {v:.2f} RL iter/s" for k, v in direct.items()))
    hr = results[head]
    peak = PEAK[head]
    # roofline of the dominant kernel family of the headline method
    t_gemm = hr["fwd_gemm_ms"] + hr["bwd_gemm_ms"]
    t_fftp = hr["fwd_fft_ms"] + hr["bwd_fft_ms"]
    tj = None
    tpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic_rl_full.json")
    if os.path.exists(tpath):
        try:
            tj = json.load(open(tpath))
        except (OSError, ValueError) as e:
            log(f"traffic summary unreadable ({e}); roofline.traffic = null")
            tj = None
    traffic_scale = None
    if tj is not None:
        # the counters belong to one plan: same precision, method and workload; a plan that
        # puts another number of planes on the FFT path streams spectra of another size, and
        # the GEMVs' counted bytes are scaled by the ratio of the spectra sizes (each GEMV
        # streams its spectra once, so its traffic is proportional to them)
        if not (tj.get("precision") == head and tj.get("method") == args.method and wl.name == tj.get("workload")):
            tj = None
        elif abs(float(tj.get("spectra_gb", -1)) - hr["spectra_gb"]) >= 0.01:
            traffic_scale = hr["spectra_gb"] / float(tj["spectra_gb"])
    if t_gemm >= t_fftp:
        gk = {"fp32": "sid::k_proj_gemm<15, 2>", "bf16x2": "sid::k_proj_gemm_sl<15, 2, 1>",
              "bf16x3": "sid::k_proj_gemm_sl<15, 2, 2>"}.get(head, "")
        kk = tj.get("kernels", {}) if tj and traffic_scale is None else {}
        traffic = (kk[gk]["read_bytes"] + kk[gk]["write_bytes"]) if gk in kk else None
        roof = {"bound": "mfma", "achieved": round(hr["gemm_achieved_tflops"], 3), "peak": peak / 1e12,
                "unit": "TFLOP/s", "frac": round(hr["gemm_frac_of_peak"], 4), "traffic": traffic,
                "kernel": "k_proj_gemm (forward and backward launches, direct planes)",
                "algorithmic_per_launch": "2*X*Y*sum(M_c^2) over direct planes"}
    elif hr.get("fft_tiled"):
        # the spectra GEMMs of the tiled FFT path, all plane groups of one forward + one backward
        # projection; the PMC traffic of the same GEMMs (counted in a run of the same plan)
        traffic = None
        if tj and tj.get("tf_gemm_bytes") and tj.get("fft_groups") == hr["fft_groups"]:
            traffic = 0.5 * float(tj["tf_gemm_bytes"])
        tb = sum(hr["tf_gemm_gb"]) * 1e9
        oc = hr.get("one_chain")
        ach = oc["tf_gemm_stream_tbps"] if oc else hr["tf_gemm_stream_tbps"]
        ms_l = 0.5 * (oc["fwd_gemm_ms"] + oc["bwd_gemm_ms"]) if oc else 0.5 * (hr["fwd_gemv_ms"] + hr["bwd_gemv_ms"])
        roof = {"bound": "hbm", "achieved": round(ach * 1e3, 1), "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": round(ach * 1e12 / HBM_PEAK, 4), "traffic": traffic,
                "traffic_source": "profiles/pmc_traffic_rl_full_bf16x2.md" if traffic else None,
                "kernel": "k_tf_gemm (spectra GEMMs of the tiled FFT path, one launch per plane group and "
                          "direction; FP16 MFMA)",
                "algorithmic_per_launch": f"per projection {0.5 * tb / 1e9:.2f} GB (forward "
                                          f"{hr['tf_gemm_gb'][0]:.2f} / backward {hr['tf_gemm_gb'][1]:.2f} GB: "
                                          f"FP16 spectra once, input spectra read, results written)",
                "ms_per_launch": round(ms_l, 3),
                "timing": "HIP events around each group's GEMM launch on its stream, one stream chain"
                          if oc else "HIP events around each group's GEMM launch on its stream",
                "achieved_two_chains": round(hr["tf_gemm_stream_tbps"] * 1e3, 1),
                "mfma_achieved_tflops": round(hr["tf_gemm_mfma_tflops"] * (ach / hr["tf_gemm_stream_tbps"]), 1),
                "mfma_frac_of_fp16_peak": round(hr["tf_gemm_mfma_frac"] * (ach / hr["tf_gemm_stream_tbps"]), 4),
                "plane_groups": hr["fft_groups"]}
    else:
        traffic = None
        if tj:
            # the GEMV instantiation of this precision's spectra (FP16 pairs in BF16X2, FP32 otherwise)
            inst = "<__half2>" if head == "bf16x2" else "<HIP_vector_type<float, 2u> >"
            kk = tj.get("kernels", {})
            names = [f"sid::k_fft_gemv_{d}{inst}" for d in ("fwd", "bwd")]
            if all(k in kk for k in names):
                traffic = 0.5 * sum(kk[k]["read_bytes"] + kk[k]["write_bytes"] for k in names)
                if traffic_scale is not None:
                    traffic *= traffic_scale
        roof = {"bound": "hbm", "achieved": round(hr["gemv_stream_tbps"] * 1e3, 1), "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": round(hr["gemv_frac_of_hbm"], 4),
                "traffic": traffic, "traffic_source": "profiles/pmc_traffic_rl_full_bf16x2.md" if traffic else None,
                "traffic_scaled_from_spectra_gb": float(tj["spectra_gb"]) if traffic and traffic_scale else None,
                "kernel": "k_fft_gemv_fwd / k_fft_gemv_bwd (spectra GEMVs of the FFT planes)",
                "algorithmic_per_launch": f"{hr['spectra_gb']:.2f} GB of spectra streamed once per launch"}
    # the MFMA half of the headline plan: the direct GEMM on its planes, beside the FFT path
    roof_mfma = None
    if hr["planes_direct"] > 0 and hr["fwd_gemm_ms"] + hr["bwd_gemm_ms"] > 0:
        prods = {"bf16x2": 2, "bf16x3": 3}.get(head, 1)
        t_launch = 0.5 * (hr["fwd_gemm_ms"] + hr["bwd_gemm_ms"]) * 1e-3
        alg = 2.0 * hr["gemm_macs_algorithmic"]
        ex = 2.0 * hr["gemm_macs_executed"] * prods
        roof_mfma = {"bound": "mfma", "achieved": round(alg / t_launch / 1e12, 2), "peak": peak / 1e12,
                     "unit": "TFLOP/s", "frac": round(alg / t_launch / peak, 4),
                     "executed_mfma_frac": round(ex / t_launch / peak, 4),
                     "executed_over_algorithmic": round(ex / prods / alg, 3),
                     "kernel": "k_proj_gemm (direct planes, forward and backward launches beside the FFT path)",
                     "algorithmic_per_launch": f"2*X*Y*sum(M_c^2) over {hr['planes_direct']} direct planes = {alg:.3e} flop",
                     "ms_per_launch": round(t_launch * 1e3, 3),
                     "plane_exec_over_alg": hr["plane_exec_over_alg"]}
    fft_planes = hr["planes_fft"]
    dtype = head
    if head == "bf16x2":
        # the reduced-precision split mode: direct planes multiply BF16 hi/lo volume values by the
        # BF16 PSF; tiled FFT planes multiply FP16 hi/lo input spectra by FP16 spectra (FP32 accumulate)
        tiled = hr.get("fft_tiled")
        if fft_planes and not hr["planes_direct"]:
            dtype = "fp16x2-split-mfma" if tiled else "fp16-spectra"
        elif fft_planes:
            dtype = "bf16x2-split-mfma+" + ("fp16x2-split-mfma" if tiled else "fp16-spectra")
        else:
            dtype = "bf16x2-split-mfma"
    out = {
        "metric": "RL iter/s @1005²×51z: BF16-split MFMA vs plain FP32; rel-L2 err vs FP64",
        "value": round(hr["iter_per_s"], 4),
        "unit": "RL iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(hr["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "plane_mix": {"direct": hr["planes_direct"], "fft": fft_planes,
                      "direct_arithmetic": {"bf16x2": "bf16 hi/lo split x bf16 PSF, fp32 accumulate",
                                            "bf16x3": "bf16 hi/lo split x bf16 hi/lo PSF, fp32 accumulate",
                                            "fp32": "fp32", "fp64": "fp64"}.get(head, head),
                      "fft_spectra": ("fp16 pairs, row-scaled" if head == "bf16x2" else "complex fp32")
                      if fft_planes else None,
                      "fft_arithmetic": ("tiled: fp16 spectra x fp16 hi/lo input spectra on MFMA, fp32 accumulate"
                                         if hr.get("fft_tiled") else "spectra GEMV, fp32 accumulate")
                      if fft_planes else None,
                      "fft_plane_groups": hr.get("fft_groups")},
        "data": "synthetic (seeded PSF, ball volume, Poisson image)",
        "config": {"workload": f"{wl.name}: {wl.X}x{wl.Y} px, Nnum={wl.n}, {wl.nz} planes z={wl.zs[0]}..{wl.zs[-1]}, "
                               f"M={wl.M}, {wl.n_balls} balls, {args.iters} ISRA iterations per step",
                   "method": head, "baseline_method": base, "projector": args.method,
                   "parallelism": "single GPU"},
        "roofline": roof,
        "roofline_mfma": roof_mfma,
        "roofline_direct_only": direct_roof if direct else None,
        "methods": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in results.items()},
        "speedup_vs_fp32": round(hr["iter_per_s"] / results[base]["iter_per_s"], 3) if base in results else None,
        "host_entry_iter_per_s": host_rate,
        "direct_only_iter_per_s": direct,
        "speedup_vs_fp32_direct_only": (round(direct[head] / direct[base], 3)
                                        if direct and head in direct and base in direct and head != base else None),
        "rel_l2_vs": ref_name,
        "rel_l2": rel,
    }
    if not args.no_library:
        out["library"] = out_lib
    if out_rows:
        out["rows"] = out_rows
    if proj_rates:
        out["projections"] = proj_rates
        out["projections"]["rl_full"] = {head: {"proj_per_s": round(hr["proj_per_s"], 1)},
                                         base: {"proj_per_s": round(results[base]["proj_per_s"], 1)}}
        if out_rows and "gradient" in out_rows:
            out["projections"]["grad"] = {"fp32": out_rows["gradient"]["projections_fp32_direct"]}
    # release the device state before the CPU leg and the exit
    htf_host = htfs[base].cpu().numpy() if not args.no_cpu else None
    psf.close()
    del Vt, yt, htfs, clean
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    if not args.no_cpu:
        log("CPU baseline (one RL iteration of the full workload, and the small workload in full) ...")
        Hd, _ = synth.make_psf(wl.n, wl.M, wl.zs, xp=torch, device=dev)
        H_host = Hd.cpu().numpy()   # (Nz, n, n, M, M) C order = MATLAB [M, M, n, n, Nz]
        del Hd
        torch.cuda.empty_cache()
        out["cpu_baseline"] = cpu_baseline(wl, H_host, htf_host)
        del H_host
        log(f"CPU baseline {out['cpu_baseline']['value']:.5f} RL iter/s on {out['cpu_baseline']['cores']} threads "
            f"({out['cpu_baseline']['cpu_model']}); rl_small {out['cpu_baseline']['rl_small']['iter_per_s']} RL iter/s")
    print(json.dumps(out), flush=True)
    return args


if __name__ == "__main__":
    if main().clean_exit:
        sys.exit(0)
    # The result is printed and the device state released; leave without the
    # interpreter's and the runtimes' teardown, so no process of the run is
    # still winding down after its last line (nothing here starts a child
    # process; the host-copy and OpenMP workers are threads of this one).
    sys.stdout.flush()
    sys.stderr.flush()
    os._exit(0)
