"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the FP64 oracle): the C oracle reproduces them (CPU), and the HIP path
matches them through the C ABI (GPU)."""
import os

import numpy as np
import pytest

from oracle import sid_oracle as O
from sid_amd import synth
from tests._helpers import c_backward, c_forward, c_isra, rel_l2

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RANDOM_CASES = ["tiny_n3_odd", "tiny_n3_even", "crop_n5"]


def load(name):
    return dict(np.load(os.path.join(HERE, name + ".npz")))


@pytest.mark.parametrize("name", RANDOM_CASES)
def test_c_oracle_reproduces_fixture(name):
    f = load(name)
    CA = f.get("CAindex")
    assert rel_l2(c_forward(f["H"], f["V"], CA), f["O"]) < 1e-13
    assert rel_l2(c_backward(f["H"], f["W"], CA), f["B"]) < 1e-13
    assert rel_l2(c_isra(f["H"], f["htf"], f["htf"], int(f["iters"]), CA), f["isra"]) < 1e-12


def test_workload_fixture_regenerates():
    """The synthetic inputs are deterministic: the oracle's forward image of the
    regenerated PSF and volume equals the stored one."""
    f = load("lf_small_n15")
    zs = tuple(int(z) for z in f["zs"])
    H, _ = synth.make_psf(int(f["n"]), int(f["M"]), zs, dtype=np.float64)
    V = synth.make_volume(int(f["X"]), int(f["X"]), len(zs), int(f["n_balls"]), 3, 3, int(f["seed"]))
    assert rel_l2(c_forward(H, V.astype(np.float64)), f["clean"]) < 1e-6
    y = synth.poisson_image(f["clean"].astype(np.float64), 1000.0, int(f["seed"]))
    yn, nrm = O.reconstruct_preprocess(y)
    assert abs(nrm - float(f["nrm"])) < 1e-9 * nrm


@pytest.mark.gpu
@pytest.mark.parametrize("name", RANDOM_CASES)
def test_gpu_matches_fixture(name):
    import sid_amd
    print("case: golden", name, flush=True)
    f = load(name)
    psf = sid_amd.PSF(f["H"].astype(np.float32), f.get("CAindex"))
    assert rel_l2(psf.forward(f["V"].astype(np.float32)), f["O"]) < 2e-6
    assert rel_l2(psf.backward(f["W"].astype(np.float32)), f["B"]) < 2e-6
    X = psf.deconv_rl(f["htf"].astype(np.float32), int(f["iters"]), f["htf"].astype(np.float32))
    assert rel_l2(X, f["isra"]) < 1e-5


@pytest.mark.gpu
def test_gpu_reconstruction_matches_workload_fixture():
    """reconstruct_S wiring (FP64 pre-processing, Htf = B(y/nrm), 8 ISRA
    iterations) on the small n = 15 workload, FP32 and BF16x2 on the device."""
    import sid_amd
    from sid_amd import recon
    print("case: golden lf_small_n15", flush=True)
    f = load("lf_small_n15")
    zs = tuple(int(z) for z in f["zs"])
    H, _ = synth.make_psf(int(f["n"]), int(f["M"]), zs)
    y = f["y"].astype(np.float64)
    for prec, tol in (("fp32", 2e-5), ("bf16x2", 1e-3), ("bf16x3", 2e-5)):
        psf = sid_amd.PSF(H, precision=prec)
        out = recon.reconstruct(psf, y, int(f["iters"]))
        ref = float(f["nrm"]) * f["isra"].astype(np.float64)
        e = rel_l2(out, ref)
        print(prec, "reconstruction rel-L2", e, flush=True)
        assert e < tol
        psf.close()
