"""The MEX gateway on the device, compiled against the test stand-in of
MATLAB's mex.h: output classes and sizes of every entry point (single for the
GPU projectors, double for backwardProjectACC_new.m:6), values against the
oracle, Xguess_history (deconvRL.m:9-11,90-92), an in-place edit of H seen by
the handle cache, explicit sidPSF handles, LS_deconv and both library
generators."""
import numpy as np
import pytest

import oracle.sid_oracle as O
from tests._helpers import c_backward, c_forward, c_isra, random_psf, rel_l2
from tests._mex import Mex, MexError

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import sid_amd
    if sid_amd.device_count() < 1:
        pytest.fail("no gfx950 device visible")
    from tests.mex_stub.build_mex import VARIANTS
    yield sid_amd
    for name in VARIANTS:                   # mexAtExit: release each gateway's cached handle
        Mex(name).exit()


def _case(seed, n=5, X=40, Y=35, M=11, nz=3):
    rng = np.random.default_rng(seed)
    return rng, random_psf(rng, M, n, nz).astype(np.float32), rng.random((X, Y, nz))


def test_projectors_classes_and_values(env):
    print("case: MEX projectors n5 X40 Y35 M11 Nz3", flush=True)
    rng, H, V = _case(1)
    H64 = H.astype(np.float64)
    cls, O_ = Mex("forwardProjectGPU")(H, V)                  # double volume: converted to single
    assert cls == "single" and O_.shape == (40, 35)
    assert rel_l2(O_, c_forward(H64, V.astype(np.float32))) < 2e-6
    img = rng.random((40, 35))
    cls, B = Mex("backwardProjectGPU_new")(H, img.astype(np.float32))
    assert cls == "single" and B.shape == (40, 35, 3)
    assert rel_l2(B, c_backward(H64, img.astype(np.float32).astype(np.float64))) < 2e-6
    CA = np.array([[2, 10], [1, 11], [3, 9]], dtype=np.float64)
    cls, Oa = Mex("forwardProjectACC")(H, V.astype(np.float32), CA)
    assert cls == "single" and rel_l2(Oa, c_forward(H64, V.astype(np.float32), CA.astype(np.int32))) < 2e-6
    cls, Ba = Mex("backwardProjectACC_new")(H, img.astype(np.float32), CA)
    assert cls == "double" and Ba.shape == (40, 35, 3)
    assert rel_l2(Ba, c_backward(H64, img.astype(np.float32).astype(np.float64), CA.astype(np.int32))) < 2e-6


def test_handle_cache_sees_an_in_place_edit(env):
    """The same H buffer with one value changed must not reuse the old weights."""
    print("case: MEX handle cache, in-place edit of H", flush=True)
    rng, H, V = _case(2)
    fwd = Mex("forwardProjectGPU")
    _, O1 = fwd(H, V.astype(np.float32))
    H[5, 5, 1, 2, 0] += 3.0                                   # one entry, same array
    _, O2 = fwd(H, V.astype(np.float32))
    assert rel_l2(O2, c_forward(H.astype(np.float64), V.astype(np.float32))) < 2e-6
    assert not np.array_equal(O1, O2)


def test_isra_history(env):
    print("case: MEX sidDeconvRL with Xguess_history", flush=True)
    rng, H, V = _case(3)
    H64 = H.astype(np.float64)
    htf = c_backward(H64, c_forward(H64, V)).astype(np.float32)
    (cx, X), (ch, hist) = Mex("sidDeconvRL")(H, htf, 3.0, htf, nlhs=2)
    assert cx == "single" and ch == "single" and hist.shape == (3, 40, 35, 3)
    assert np.array_equal(hist[2], X)
    Xo = htf.astype(np.float64)
    for i in range(3):
        Xo = c_isra(H64, htf.astype(np.float64), Xo, 1)
        assert rel_l2(hist[i], Xo) < 1e-5
    cls, X1 = Mex("sidDeconvRL")(H, htf, 3.0, htf)             # one output: no history
    assert np.array_equal(X1, X)


def test_explicit_handles_and_ls_deconv(env):
    sid_amd = env
    print("case: MEX sidPSF handles and sidLSDeconv", flush=True)
    rng, H, V = _case(4)
    psf_mex = Mex("sidPSF")
    cls, h = psf_mex("create", H)
    assert cls == "uint64" and h.shape == (1, 1)
    _, O_ = Mex("forwardProjectGPU")(h, V.astype(np.float32))
    assert rel_l2(O_, c_forward(H.astype(np.float64), V.astype(np.float32))) < 2e-6
    psf_mex("precision", h, "bf16x2", nlhs=0)
    _, Ob = Mex("forwardProjectGPU")(h, V.astype(np.float32))
    assert 1e-7 < rel_l2(Ob, O_) < 3e-3
    psf_mex("precision", h, "fp32", nlhs=0)
    df = c_backward(H.astype(np.float64), c_forward(H.astype(np.float64), V)).astype(np.float32)
    opts = {"max_iter": 3.0, "lamb_L1": 0.01, "lamb_TV_L2": 0.02}
    cls, X = Mex("sidLSDeconv")(h, df, opts)
    psf = sid_amd.PSF(H)
    want = psf.ls_deconv(df, {"max_iter": 3, "lamb_L1": 0.01, "lamb_TV_L2": (0.02, 0.02, 0.02)})
    psf.close()
    # the two handles may autotune different plane splits; LS_deconv's bar is 1e-4 at 3 iterations
    assert cls == "single" and rel_l2(X, want) < 1e-4
    psf_mex("destroy", h, nlhs=0)
    with pytest.raises(MexError) as e:
        Mex("forwardProjectGPU")(h, V.astype(np.float32))
    assert e.value.ident == "sid:arg"


def test_handle_kernel_wrapping(env):
    """sidPSF('kernel', h, k): the handle's projectors and ISRA become
    F(convn(X, k, 'same')) / convn(B(y), k, 'same') (reconstruct_S.m:122-128,
    184-190); k = [] restores the plain projectors."""
    print("case: MEX sidPSF kernel", flush=True)
    rng, H, V = _case(6)
    H64 = H.astype(np.float64)
    k = rng.random((5, 3, 3))
    k = (k / np.linalg.norm(k)).astype(np.float32)
    psf_mex = Mex("sidPSF")
    _, h = psf_mex("create", H)
    psf_mex("kernel", h, k, nlhs=0)
    Vs = V.astype(np.float32)
    _, O_ = Mex("forwardProjectGPU")(h, Vs)
    want = c_forward(H64, O.convn_same(Vs.astype(np.float64), k.astype(np.float64)))
    assert rel_l2(O_, want) < 5e-6
    img = rng.random((40, 35)).astype(np.float32)
    _, B_ = Mex("backwardProjectGPU_new")(h, img)
    assert rel_l2(B_, O.convn_same(c_backward(H64, img), k.astype(np.float64))) < 5e-6
    psf_mex("kernel", h, np.zeros((0, 0)), nlhs=0)
    _, O2 = Mex("forwardProjectGPU")(h, Vs)
    assert rel_l2(O2, c_forward(H64, Vs)) < 2e-6
    psf_mex("destroy", h, nlhs=0)


def test_library_gpu_gateway(env):
    print("case: MEX generate_LFM_library_GPU", flush=True)
    rng = np.random.default_rng(5)
    n, M, X, Y, nz = 5, 21, 45, 40, 6
    H = random_psf(rng, M, n, nz).astype(np.float32)
    recon = [rng.random((X, Y, nz)).astype(np.float32) for _ in range(3)]
    centers = np.array([[20.4, 18.5, 3.2], [2.0, 39.6, 5.8], [44.5, 1.2, 1.0]])
    neur_id = np.array([[1.0, 0.0, 3.0], [0.0, 2.0, 0.0], [1.0, 1.0, 1.0]])
    opts = {"image_size": np.array([X, Y], dtype=np.float64), "axial": 2.0, "neur_rad": 3.0}
    cls, fm = Mex("generate_LFM_library_GPU")(recon, centers, neur_id, {"H": H}, opts)
    assert cls == "single" and fm.shape == (3, X * Y)
    for s in range(3):
        ids = list(np.flatnonzero(neur_id[s]))
        want, _ = O.library_gpu_entry([r.astype(np.float64) for r in recon], ids, centers[s], 3, 2.0,
                                      H.astype(np.float64))
        assert rel_l2(fm[s], want.ravel(order="F")) < 2e-6


def test_library_gpu_gateway_with_kernel(env):
    """options.ker_param: the gateway calls back MATLAB's generate_kernel
    (here the test's stand-in, sid_amd.recon.generate_kernel) and convolves each
    seed's volume before the projection (generate_LFM_library_GPU.m:50-56,95-101)."""
    from sid_amd import recon as R
    from tests._mex import MATLAB_FUNCTIONS
    print("case: MEX generate_LFM_library_GPU with ker_param", flush=True)
    calls = []

    def gen(shape, param):
        calls.append((shape, np.asarray(param).ravel().tolist()))
        return R.generate_kernel(shape, np.asarray(param).ravel())

    MATLAB_FUNCTIONS["generate_kernel"] = gen
    try:
        rng = np.random.default_rng(7)
        n, M, X, Y, nz = 5, 21, 45, 40, 6
        H = random_psf(rng, M, n, nz).astype(np.float32)
        recon = [rng.random((X, Y, nz)).astype(np.float32) for _ in range(2)]
        centers = np.array([[2.4, 20.5, 1.2], [44.5, 38.6, 6.0], [22.5, 17.4, 3.5]])
        neur_id = np.array([[1.0, 1.0], [0.0, 1.0], [1.0, 0.0]])
        opts = {"image_size": np.array([X, Y], dtype=np.float64), "axial": 2.0, "neur_rad": 3.0,
                "ker_shape": "ball", "ker_param": np.array([2.0, 1.0])}
        cls, fm = Mex("generate_LFM_library_GPU")(recon, centers, neur_id, {"H": H}, opts)
    finally:
        MATLAB_FUNCTIONS.pop("generate_kernel", None)
    assert calls == [("ball", [2.0, 1.0])]
    assert cls == "single" and fm.shape == (3, X * Y)
    kern = R.generate_kernel("ball", [2.0, 1.0]).astype(np.float64)
    for s in range(3):
        ids = list(np.flatnonzero(neur_id[s]))
        want, _ = O.library_gpu_entry([r.astype(np.float64) for r in recon], ids, centers[s], 3, 2.0,
                                      H.astype(np.float64), kernel=kern)
        assert rel_l2(fm[s], want.ravel(order="F")) < 2e-6


def test_library_cpu_gateway(env):
    print("case: MEX generate_LFM_library_CPU (sparse)", flush=True)
    rng = np.random.default_rng(6)
    n, M, m = 5, 21, np.array([50.0, 44.0, 5.0])
    H = random_psf(rng, M, n, 5).astype(np.float32)
    centers = np.array([[25.2, 20.5, 3.0], [1.6, 43.5, 4.6], [-30.0, 4.0, 2.0]])
    kind, fm = Mex("generate_LFM_library_CPU")(centers, {"H": H, "Nnum": float(n)}, 2.0, np.array([1.0, 1.0, 1.5]), m)
    assert kind == "sparse" and fm.shape == (3, 50 * 44)
    fm = fm.tocsr()
    for s in range(3):
        row = fm[s].toarray().ravel()
        try:
            Q, _ = O.library_cpu_entry(centers[s], 2, (1.0, 1.0, 1.5), (50, 44, 5), H.astype(np.float64), n)
        except ValueError:
            assert fm[s].nnz == 0
            continue
        assert np.array_equal(np.flatnonzero(row), np.flatnonzero(Q))
        assert rel_l2(row, Q) < 2e-6
