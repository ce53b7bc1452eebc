"""Writes the golden fixtures in this folder from the FP64 oracle.

The reference ships no test vectors (it has no tests), so these fixtures are
the oracle's FP64 outputs on seeded inputs: they pin the oracle against
regressions and give the GPU tests known answers.  Run from the repo root:

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import sid_oracle as O  # noqa: E402
from sid_amd import synth  # noqa: E402


def _case_random(name, seed, n, X, Y, M, nz, crop=False, iters=5):
    rng = np.random.default_rng(seed)
    H = rng.random((M, M, n, n, nz))
    V = rng.random((X, Y, nz))
    W = rng.random((X, Y))
    CA = None
    if crop:
        m = M // 2
        CA = np.array([[m + 1 - r, m + 1 + r] for r in range(1, nz + 1)], dtype=np.int32)
        F = lambda v: O.forward_project_acc(H, v, CA)
        B = lambda o: O.backward_project_acc(H, o, CA)
    else:
        F = lambda v: O.forward_project(H, v)
        B = lambda o: O.backward_project(H, o)
    out = dict(H=H, V=V, W=W, O=F(V), B=B(W))
    htf = B(F(V))
    out["htf"] = htf
    out["isra"] = O.deconv_rl(F, B, htf, iters, htf)
    out["iters"] = np.int64(iters)
    if CA is not None:
        out["CAindex"] = CA
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)


def _case_workload(name, n, X, M, zs, n_balls, seed, iters=8):
    """Small light-field workload with the synthetic PSF family: H and V are
    regenerated from the seed by sid_amd.synth; only outputs are stored."""
    H, CA = synth.make_psf(n, M, zs, dtype=np.float64)
    V = synth.make_volume(X, X, len(zs), n_balls, 3, 3, seed).astype(np.float64)
    F = lambda v: O.forward_project(H, v)
    B = lambda o: O.backward_project(H, o)
    clean = F(V)
    y = synth.poisson_image(clean, 1000.0, seed)
    yn, nrm = O.reconstruct_preprocess(y)
    htf = B(yn)
    X8 = O.deconv_rl(F, B, htf, iters, htf)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), clean=clean.astype(np.float32), y=y.astype(np.float32),
                        nrm=np.float64(nrm), htf=htf.astype(np.float32), isra=X8.astype(np.float32),
                        iters=np.int64(iters), n=np.int64(n), X=np.int64(X), M=np.int64(M),
                        zs=np.asarray(zs, dtype=np.int64), n_balls=np.int64(n_balls), seed=np.int64(seed))


if __name__ == "__main__":
    _case_random("tiny_n3_odd", 11, 3, 9, 12, 7, 3)
    _case_random("tiny_n3_even", 12, 3, 10, 8, 6, 2)
    _case_random("crop_n5", 13, 5, 25, 30, 11, 3, crop=True)
    _case_workload("lf_small_n15", 15, 105, 105, (-1, 0, 1), 10, 21)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))
