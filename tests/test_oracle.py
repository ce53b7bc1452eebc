"""Pins the oracle (CPU, no GPU): library routines, brute force, closed forms
and invariants, chosen so that a dropped term, a wrong sign or offset, or a
transposed operand fails at least one check."""
import math

import numpy as np
import pytest
from scipy.signal import convolve2d

from oracle import sid_oracle as O
from sid_amd import synth
from tests._helpers import c_backward, c_forward, c_isra, rel_l2


# ---------------------------------------------------------------- conv2 'same'

@pytest.mark.parametrize("shape_x,shape_k", [((13, 11), (7, 5)), ((9, 9), (9, 9)), ((5, 17), (3, 11)), ((12, 10), (1, 1))])
def test_conv2fft_equals_scipy_same_for_odd_kernels(shape_x, shape_k):
    rng = np.random.default_rng(0)
    x = rng.standard_normal(shape_x)
    k = rng.standard_normal(shape_k)
    # for odd kernels MATLAB's and scipy's 'same' windows coincide
    assert np.abs(O.conv2fft_same(x, k) - convolve2d(x, k, mode="same")).max() < 1e-12


def test_conv2_same_matlab_window_for_even_kernels():
    # MATLAB documentation example: conv([1 2 3], [1 1], 'same') = [3 5 3]
    x = np.array([[1.0, 2.0, 3.0]])
    k = np.array([[1.0, 1.0]])
    assert np.allclose(O.conv2_same_direct(x, k), [[3.0, 5.0, 3.0]])
    assert np.allclose(O.conv2fft_same(x, k), [[3.0, 5.0, 3.0]])
    # general even kernels: the 'full' convolution cropped at floor(M/2)
    rng = np.random.default_rng(1)
    x = rng.standard_normal((11, 9))
    k = rng.standard_normal((6, 4))
    full = convolve2d(x, k, mode="full")
    want = full[3:3 + 11, 2:2 + 9]
    assert np.abs(O.conv2fft_same(x, k) - want).max() < 1e-12
    assert np.abs(O.conv2_same_direct(x, k) - want).max() < 1e-12


def test_fft_pad_rule_values():
    # forwardProjectGPU.m:12-13: min(2^ceil(log2(s)), 256*ceil(s/256)), s = X + floor(M/2)
    assert O.fft_pad_size(255, 105) == 512          # s = 307
    assert O.fft_pad_size(1005, 345) == 1280        # s = 1177
    assert O.fft_pad_size(495, 225) == 768          # s = 607
    assert O.fft_pad_size(100, 9) == 128            # s = 104: power of two wins
    for X, M in [(255, 105), (1005, 345), (495, 225), (37, 11)]:
        assert O.fft_pad_size(X, M) >= X + M // 2   # no wrap-around inside the 'same' window


def test_conv2fft_no_aliasing_at_pad_size():
    rng = np.random.default_rng(2)
    x = rng.random((40, 33))
    k = rng.random((21, 21))
    assert np.abs(O.conv2fft_same(x, k) - O.conv2_same_direct(x, k)).max() < 1e-12


# ---------------------------------------------------------------- projectors

def _tiny(rng, n=3, X=9, Y=12, M=7, nz=3):
    H = rng.random((M, M, n, n, nz))
    V = rng.random((X, Y, nz))
    return H, V


def test_forward_equals_dense_brute_force():
    rng = np.random.default_rng(3)
    H, V = _tiny(rng)
    A = O.dense_operator(H, 9, 12)
    ref = (A @ V.ravel(order="F")).reshape((9, 12), order="F")
    assert np.abs(O.forward_project(H, V) - ref).max() < 1e-12
    assert np.abs(c_forward(H, V) - ref).max() < 1e-12


def test_backward_is_dense_transpose():
    rng = np.random.default_rng(4)
    H, _ = _tiny(rng)
    w = rng.random((9, 12))
    A = O.dense_operator(H, 9, 12)
    ref = (A.T @ w.ravel(order="F")).reshape((9, 12, 3), order="F")
    assert np.abs(O.backward_project(H, w) - ref).max() < 1e-12
    assert np.abs(c_backward(H, w) - ref).max() < 1e-12


def test_even_kernel_forward_and_adjoint():
    rng = np.random.default_rng(5)
    H, V = _tiny(rng, M=6)
    A = O.dense_operator(H, 9, 12)
    f = O.forward_project(H, V)
    assert np.abs(f - (A @ V.ravel(order="F")).reshape((9, 12), order="F")).max() < 1e-12
    # the reference backward rotates the even kernel by 180 degrees; that is
    # the exact transpose of the forward only for odd M.  The oracle follows
    # the reference (rot180 + conv2 'same'), checked against the defining sum.
    w = rng.random((9, 12))
    B = O.backward_project(H, w)
    for c in range(3):
        for a in range(3):
            for b in range(3):
                t = O.conv2_same_direct(w, H[::-1, ::-1, a, b, c])
                assert np.abs(B[a::3, b::3, c] - t[a::3, b::3]).max() < 1e-12
    assert np.abs(c_backward(H, w) - B).max() < 1e-12


def test_point_source_closed_form():
    """A unit voxel at q gives H_{a(q),b(q),c} centred at q (conv2 'same')."""
    n, M = 5, 11
    H, _ = synth.make_psf(n, M, (-1, 0, 2), dtype=np.float64)
    X = Y = 40
    q = (17, 23)
    for c in range(3):
        V = np.zeros((X, Y, 3))
        V[q[0], q[1], c] = 1.0
        img = O.forward_project(H, V)
        a, b = q[0] % n, q[1] % n
        want = np.zeros((X, Y))
        m = M // 2
        want[q[0] - m:q[0] + m + 1, q[1] - m:q[1] + m + 1] = H[:, :, a, b, c]
        assert np.abs(img - want).max() < 1e-13
        assert np.abs(c_forward(H, V) - want).max() < 1e-13
    # the brute-force matrix's column of voxel (q, c) is the same closed form
    A = O.dense_operator(H[:, :, :, :, :1], 20, 20, 1)
    q = (7, 12)
    want = np.zeros((20, 20))
    want[q[0] - M // 2:q[0] + M // 2 + 1, q[1] - M // 2:q[1] + M // 2 + 1] = H[:, :, q[0] % n, q[1] % n, 0]
    assert np.abs(A[:, q[0] + 20 * q[1]].reshape(20, 20, order="F") - want).max() == 0.0


def test_linearity_and_flux_conservation():
    n, M = 5, 15
    H, _ = synth.make_psf(n, M, (-2, 0, 2), dtype=np.float64)
    rng = np.random.default_rng(6)
    X = Y = 50
    V1 = np.zeros((X, Y, 3))
    V1[15:35, 15:35, :] = rng.random((20, 20, 3))   # interior: the whole footprint lands on the sensor
    V2 = rng.random((X, Y, 3))
    f1, f2 = O.forward_project(H, V1), O.forward_project(H, V2)
    assert np.abs(O.forward_project(H, 2.0 * V1 - 3.0 * V2) - (2.0 * f1 - 3.0 * f2)).max() < 1e-11
    # unit-sum kernels conserve flux for interior sources
    assert abs(f1.sum() - V1.sum()) < 1e-10 * V1.sum()


def test_lenslet_shift_equivariance():
    """Shifting the volume by a whole lenslet shifts the image by n pixels."""
    n, M = 3, 9
    rng = np.random.default_rng(7)
    H = rng.random((M, M, n, n, 2))
    V = np.zeros((30, 30, 2))
    V[9:18, 9:18, :] = rng.random((9, 9, 2))
    Vs = np.roll(V, shift=(n, 2 * n), axis=(0, 1))
    f, fs = O.forward_project(H, V), O.forward_project(H, Vs)
    assert np.abs(np.roll(f, shift=(n, 2 * n), axis=(0, 1)) - fs).max() < 1e-12


def test_adjoint_identity_random():
    rng = np.random.default_rng(8)
    H = rng.random((13, 13, 5, 5, 3))
    v = rng.random((33, 28, 3))
    w = rng.random((33, 28))
    lhs = np.sum(c_forward(H, v) * w)
    rhs = np.sum(v * c_backward(H, w))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_acc_crop_semantics():
    """forwardProjectACC crops each kernel to CAindex and uses the crop's own
    'same' offset; with zeros outside the crop it equals the full-kernel path."""
    rng = np.random.default_rng(9)
    M, n = 11, 3
    H = rng.random((M, M, n, n, 2))
    CA = np.array([[3, 9], [2, 10]], dtype=np.int32)    # symmetric crops of odd size
    Hz = np.zeros_like(H)
    for c in range(2):
        lo, hi = CA[c, 0] - 1, CA[c, 1]
        Hz[lo:hi, lo:hi, :, :, c] = H[lo:hi, lo:hi, :, :, c]
    V = rng.random((15, 12, 2))
    fa = O.forward_project_acc(H, V, CA)
    assert np.abs(fa - O.forward_project(Hz, V)).max() < 1e-12
    assert np.abs(c_forward(H, V, CA) - fa).max() < 1e-12
    w = rng.random((15, 12))
    ba = O.backward_project_acc(H, w, CA)
    assert np.abs(ba - O.backward_project(Hz, w)).max() < 1e-12
    assert np.abs(c_backward(H, w, CA) - ba).max() < 1e-12
    # an asymmetric / even crop uses floor(Mc/2) of the crop
    CA2 = np.array([[2, 7], [4, 11]], dtype=np.int32)
    fa2 = O.forward_project_acc(H, V, CA2)
    want = np.zeros((15, 12))
    for a in range(n):
        for b in range(n):
            for c in range(2):
                t = np.zeros((15, 12))
                t[a::n, b::n] = V[a::n, b::n, c]
                K = H[CA2[c, 0] - 1:CA2[c, 1], CA2[c, 0] - 1:CA2[c, 1], a, b, c]
                want += O.conv2_same_direct(t, K)
    assert np.abs(fa2 - want).max() < 1e-12
    assert np.abs(c_forward(H, V, CA2) - want).max() < 1e-12


def test_c_and_numpy_oracles_agree_at_small_workload():
    n, M = 15, 45
    H, CA = synth.make_psf(n, M, (-2, 0, 2), dtype=np.float64)
    rng = np.random.default_rng(10)
    V = rng.random((75, 60, 3))
    assert rel_l2(c_forward(H, V), O.forward_project(H, V)) < 1e-13
    w = rng.random((75, 60))
    assert rel_l2(c_backward(H, w), O.backward_project(H, w)) < 1e-13


# ---------------------------------------------------------------- ISRA

def test_isra_equals_dense_isra_and_guards():
    rng = np.random.default_rng(11)
    H, V = _tiny(rng, nz=2)
    A = O.dense_operator(H, 9, 12)
    y = A @ V.ravel(order="F")
    htf = A.T @ y
    x = htf.copy()
    for _ in range(5):
        x = x * htf / (A.T @ (A @ x))
    F = lambda v: O.forward_project(H, v)
    B = lambda o: O.backward_project(H, o)
    htf3 = htf.reshape((9, 12, 2), order="F")
    X = O.deconv_rl(F, B, htf3, 5, htf3)
    assert np.abs(X.ravel(order="F") - x).max() < 1e-10 * x.max()
    assert np.abs(c_isra(H, htf3, htf3, 5) - X).max() < 1e-10 * x.max()


def test_isra_fixed_point_and_monotone_decrease():
    rng = np.random.default_rng(12)
    H = rng.random((9, 9, 3, 3, 2))
    Vt = rng.random((12, 15, 2)) + 0.1
    F = lambda v: O.forward_project(H, v)
    B = lambda o: O.backward_project(H, o)
    y = F(Vt)
    htf = B(y)
    # the true volume is a fixed point
    X = O.deconv_rl(F, B, htf, 3, Vt)
    assert np.abs(X - Vt).max() < 1e-10
    # from Htf the least-squares residual never increases (ISRA, De Pierro 1987)
    X = htf.copy()
    res = [np.linalg.norm(y - F(X))]
    for _ in range(6):
        X = O.deconv_rl(F, B, htf, 1, X)
        res.append(np.linalg.norm(y - F(X)))
    assert all(res[i + 1] <= res[i] * (1 + 1e-12) for i in range(len(res) - 1))


def test_isra_zero_guards():
    """deconvRL.m:85-87: x/0 -> Inf -> 0; 0/0 -> NaN -> X set to 0."""
    F = lambda v: np.zeros((2, 2))
    B = lambda o: np.zeros((2, 2, 1))
    htf = np.array([[1.0, 0.0], [2.0, 0.0]])[:, :, None]
    X = O.deconv_rl(F, B, htf, 1, np.ones((2, 2, 1)))
    assert np.all(X == 0.0)


# ---------------------------------------------------------------- reconstruct_S pieces

def test_quantile_hazen_hand_worked():
    x = np.array([3.0, 1.0, 4.0, 1.0, 5.0])   # sorted 1 1 3 4 5, p_k = (k - 0.5)/5
    assert O.quantile_matlab(x, 0.5) == 3.0
    assert O.quantile_matlab(x, 0.1) == 1.0                   # p_1
    assert O.quantile_matlab(x, 0.05) == 1.0                  # clamped below p_1
    assert O.quantile_matlab(x, 0.99) == 5.0                  # clamped above p_5
    assert abs(O.quantile_matlab(x, 0.8) - 4.5) < 1e-15       # halfway between p_4 = 0.7 and p_5 = 0.9
    xs = np.random.default_rng(0).random(101)
    for q in (0.2, 0.8, 0.95):
        assert abs(O.quantile_matlab(xs, q) - np.quantile(xs, q, method="hazen")) < 1e-15


def test_preprocess_hand_worked():
    """reconstruct_S.m:131-134,168-170,191 on values worked by hand.

    img(:) = 1..12: the Hazen median of 1..12 is 6.5; after the threshold the
    values 7..12 become 0.5..5.5, divided by 5.5; the norm is sqrt(71.5)/5.5."""
    img = np.arange(1.0, 13.0).reshape(3, 4, order="F")
    y, nrm = O.reconstruct_preprocess(img, None, 0.5)
    assert abs(nrm - math.sqrt(71.5) / 5.5) < 1e-15
    want = np.maximum(img - 6.5, 0) / math.sqrt(71.5)
    assert np.abs(y - want).max() < 1e-15
    # with a microlens mask the threshold comes from the unmasked pixels only:
    # values 1..4 (mask == 0) have Hazen median 2.5
    mask = np.ones(12)
    mask[:4] = 0
    y, nrm = O.reconstruct_preprocess(img, mask.reshape(3, 4, order="F"), 0.5)
    z = np.maximum(np.arange(1.0, 13.0) - 2.5, 0) / 9.5
    assert abs(nrm - math.sqrt(float((z ** 2).sum()))) < 1e-15
    assert np.abs(y.ravel(order="F") - z / math.sqrt(float((z ** 2).sum()))).max() < 1e-15


def test_reconstruct_isra_is_dense_isra_of_the_preprocessed_image():
    """reconstruct_S.m:168-205 (ISRA): recon = nrm * deconvRL(F, B, B(y/nrm), it, B(y/nrm)),
    recomposed here from the dense matrix and numpy's Hazen quantile."""
    n, M, X, Y, nz = 3, 5, 6, 6, 2
    rng = np.random.default_rng(21)
    H = rng.random((M, M, n, n, nz))
    img = rng.random((X, Y)) * 10
    A = O.dense_operator(H, X, Y, nz)
    thr = np.quantile(img.ravel(), 0.8, method="hazen")
    z = np.maximum(img - thr, 0)
    z = z / z.max()
    nrm = np.linalg.norm(z)
    yv = (z / nrm).ravel(order="F")
    htf = A.T @ yv
    x = htf.copy()
    for _ in range(5):
        x = x * (htf / (A.T @ (A @ x)))
    got = O.reconstruct_isra(H, img, max_iter=5)
    want = (nrm * x).reshape(X, Y, nz, order="F")
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_matlab_colon_counts_and_values():
    """MATLAB a:b has fix(b - a) + 1 elements starting at a (colon documentation)."""
    assert list(O.matlab_colon(-2, 2)) == [-2, -1, 0, 1, 2]
    v = O.matlab_colon(-4 / 3, 4 / 3)              # -rr:rr with rr = 4/3: 3 elements, not 4
    assert v.size == 3 and abs(v[-1] - 2 / 3) < 1e-15
    assert O.matlab_colon(-1.5, 1.5).size == 4       # 2 rr = 3 exactly
    assert O.matlab_colon(-1.2, 1.2).size == 3
    assert O.matlab_colon(0.0, 0.3 * 3 / 0.3).size == 4   # round-off in (b - a) is tolerated
    assert O.matlab_colon(1, 0).size == 0


def test_matlab_ind2sub3_hand_worked():
    # siz [3 3 x]: index 14 -> (2, 2, 2); index 5 -> (2, 2, 1); the third size does not enter
    for siz in ((3, 3, 3), (3, 3, 2.5)):
        i, j, k = O.matlab_ind2sub3(siz, [5, 14, 23, 1, 27])
        assert list(i) == [2, 2, 2, 1, 3] and list(j) == [2, 2, 2, 1, 3] and list(k) == [1, 2, 3, 1, 3]


def test_ball_offsets_hand_counted():
    """generate_LFM_library_CPU.m:34-39 on balls counted by hand."""
    BWW, W = O.ball_offsets(1, 1.0)
    # X^2 + Y^2 + Z^2 <= 1: the centre and its 6 neighbours, in find() order
    assert [tuple(int(v) for v in r) for r in BWW] == [
        (2, 2, 1), (2, 1, 2), (1, 2, 2), (2, 2, 2), (3, 2, 2), (2, 3, 2), (2, 2, 3)]
    # r = 4, rr = 4/3 (dim(3) = 3): -rr:rr = [-4/3, -1/3, 2/3] -> 3 planes; Z*r/rr = -4, -1, 2
    # -> 1 + #{x^2+y^2 <= 15} + #{x^2+y^2 <= 11} = 1 + 45 + 37 voxels
    BWW, W = O.ball_offsets(4, 4 / 3)
    assert W.shape == (9, 9, 3)
    assert [int((BWW[:, 2] == k).sum()) for k in (1, 2, 3)] == [1, 45, 37]
    # r = 2, rr = 1.5: 4 planes with Z*r/rr = -2, -2/3, 2/3, 2 -> 1, 9, 9, 1
    BWW, W = O.ball_offsets(2, 1.5)
    assert W.shape == (5, 5, 4)
    assert [int((BWW[:, 2] == k).sum()) for k in (1, 2, 3, 4)] == [1, 9, 9, 1]
    # (i, j) = (row of Y, column of X): the voxel list is symmetric, and its mean is the centre
    assert np.allclose(BWW[:, :2].mean(axis=0), [3, 3])


def test_library_cpu_voxels_hand_worked():
    """generate_LFM_library_CPU.m:43-50: round(BWW - (r+1, r+1, rr+1) + centre), clipped."""
    BWW, _ = O.ball_offsets(1, 1.0)
    B, B_ = O.library_cpu_voxels((1.4, 10.0, 2.0), 1, 1.0, BWW, (20, 20), 3)
    # x = round(i - 0.6): i = 1 -> 0 (dropped), 2 -> 1, 3 -> 2; y = j + 8; z = k
    assert [tuple(r) for r in B] == [(1, 10, 1), (1, 9, 2), (1, 10, 2), (2, 10, 2), (1, 11, 2), (1, 10, 3)]
    assert [tuple(r) for r in B_] == [(2, 2, 1), (2, 1, 2), (2, 2, 2), (3, 2, 2), (2, 3, 2), (2, 2, 3)]
    # rr = 1.5: z = round(k - 2.5 + 5.2) = k + 3 for k = 1..4; planes above size(H,5) = 6 dropped
    BWW, _ = O.ball_offsets(2, 1.5)
    B, _ = O.library_cpu_voxels((10.0, 10.0, 5.2), 2, 1.5, BWW, (20, 20), 6)
    assert sorted(set(int(z) for z in B[:, 2])) == [4, 5, 6]
    assert B.shape[0] == 19


def _full_projection_of(H, vox, vals, X, Y):
    V = np.zeros((X, Y, H.shape[4]))
    for (x, y, z), v in zip(vox, vals):
        V[x - 1, y - 1, z - 1] = v
    return O.forward_project(H, V)


@pytest.mark.parametrize("center", [(12.3, 14.6, 2.4), (2.2, 13.0, 1.0), (23.6, 24.5, 3.0)])
def test_library_cpu_entry_is_the_normalised_patch_of_the_full_projection(center):
    """generate_LFM_library_CPU.m:51-53: Q = patched projection / its norm.  On the
    pasted rectangle it equals the full-sensor projection of the same voxels
    (odd n and M), zero elsewhere; the centres include clipped borders."""
    n, M, m = 3, 7, (24, 25, 3)
    rng = np.random.default_rng(5)
    H = rng.random((M, M, n, n, 3))
    Q, info = O.library_cpu_entry(center, 2, (1.0, 1.0, 1.5), m, H, n)
    assert abs(np.linalg.norm(Q) - 1.0) < 1e-14
    full = _full_projection_of(H, info["voxels"], np.ones(len(info["voxels"])), m[0], m[1])
    win = np.zeros_like(full)
    x0, x1, y0, y1 = info["x0"], info["x1"], info["y0"], info["y1"]
    win[x0 - 1:x1, y0 - 1:y1] = full[x0 - 1:x1, y0 - 1:y1]
    want = win.ravel(order="F") / np.linalg.norm(win)
    assert np.abs(Q - want).max() < 1e-12


def test_library_cpu_kernel_path_and_sparse_rows():
    """generate_LFM_library_CPU.m:29-32 (kernel given: W = kernel, voxels = find(W),
    values W) and :54-70 (rows I = centre, J = find(Q), S = Q(J))."""
    n, M, m = 3, 5, (18, 18, 4)
    rng = np.random.default_rng(8)
    H = rng.random((M, M, n, n, 4))
    ker = rng.random((3, 3, 3)) + 0.1
    ker[0, 0, :] = 0
    centers = np.array([[9.0, 9.0, 2.0], [4.4, 15.5, 3.0]])
    row_ptr, cols, vals = O.library_cpu(centers, None, None, m, H, n, kernel=ker)
    for s, c in enumerate(centers):
        Q, info = O.library_cpu_entry(c, None, None, m, H, n, kernel=ker)
        J = np.flatnonzero(Q)
        assert np.array_equal(cols[row_ptr[s]:row_ptr[s + 1]], J)
        assert np.array_equal(vals[row_ptr[s]:row_ptr[s + 1]], Q[J])
        # r = rr = 1: voxel (i, j, k) goes to round((i, j, k) - 2 + centre) with value ker(i, j, k)
        vox, val = [], []
        for k in range(3):
            for j in range(3):
                for i in range(3):
                    if ker[i, j, k] == 0:
                        continue
                    p = O.matlab_round(np.array([i + 1, j + 1, k + 1]) - 2.0 + c).astype(int)
                    if 0 < p[0] <= m[0] and 0 < p[1] <= m[1] and 0 < p[2] <= 4:
                        vox.append(p)
                        val.append(ker[i, j, k])
        assert len(vox) == len(info["voxels"])
        full = _full_projection_of(H, vox, val, m[0], m[1])
        x0, x1, y0, y1 = info["x0"], info["x1"], info["y0"], info["y1"]
        win = np.zeros_like(full)
        win[x0 - 1:x1, y0 - 1:y1] = full[x0 - 1:x1, y0 - 1:y1]
        assert np.abs(Q - win.ravel(order="F") / np.linalg.norm(win)).max() < 1e-12


def test_library_cpu_entry_outside_the_volume_raises():
    H = np.ones((5, 5, 3, 3, 2))
    with pytest.raises(ValueError):
        O.library_cpu_entry((-10.0, 5.0, 1.0), 1, (1, 1, 1), (12, 12, 2), H, 3)


def test_library_gpu_entry_is_box_average_on_its_planes():
    """generate_LFM_library_GPU.m:76-95: the clipped box of each listed component,
    summed and divided by their count, projected with H(:,:,:,:,z0:z1) - equal to
    the full-H projection of that box placed at planes z0..z1 of a zero volume."""
    n, M, X, Y, Nz = 3, 5, 15, 14, 5
    rng = np.random.default_rng(3)
    H = rng.random((M, M, n, n, Nz))
    recon = [rng.random((X, Y, Nz)) for _ in range(3)]
    img, p = O.library_gpu_entry(recon, [0, 2], (13.6, 3.2, 4.4), 3, 2.0, H)
    # round -> (14, 3, 4); dz = ceil(3/2) = 2 -> planes 2..5 (clipped at Nz); box x 11..15, y 1..6
    assert (p["x0"], p["x1"], p["y0"], p["y1"], p["z0"], p["z1"]) == (11, 15, 1, 6, 2, 5)
    V = np.zeros((X, Y, Nz))
    V[10:15, 0:6, 1:5] = 0.5 * (recon[0][10:15, 0:6, 1:5] + recon[2][10:15, 0:6, 1:5])
    assert np.abs(img - O.forward_project(H, V)).max() < 1e-12 * np.abs(img).max()


def test_library_gpu_entry_with_kernel():
    """generate_LFM_library_GPU.m:95-101 with a kernel: the seed volume (zero
    outside the box, image_size x its planes) convolved by convn 'same' before
    the projection.  Pinned by scipy's full convolution cropped at MATLAB's
    'same' window (floor(size(k)/2)), by a delta kernel (the kernel-free entry),
    and by the z crop: with a kernel that only shifts content down in z the
    last plane of the sub-volume receives nothing."""
    from scipy.signal import convolve
    n, M, X, Y, Nz = 3, 5, 16, 14, 6
    rng = np.random.default_rng(4)
    H = rng.random((M, M, n, n, Nz))
    recon = [rng.random((X, Y, Nz)) for _ in range(2)]
    centre = (2.4, 7.6, 5.2)   # box touches the x = 1 edge; planes 3..6 touch the last plane
    for k in (rng.random((5, 3, 3)), rng.random((4, 3, 2))):
        img, p = O.library_gpu_entry(recon, [0, 1], centre, 3, 1.5, H, kernel=k)
        V = np.zeros((X, Y, p["z1"] - p["z0"] + 1))
        sx, sy, sz = slice(p["x0"] - 1, p["x1"]), slice(p["y0"] - 1, p["y1"]), slice(p["z0"] - 1, p["z1"])
        V[sx, sy, :] = 0.5 * (recon[0][sx, sy, sz] + recon[1][sx, sy, sz])
        full = convolve(V, k, mode="full")
        o = [s // 2 for s in k.shape]
        Vc = full[o[0]:o[0] + X, o[1]:o[1] + Y, o[2]:o[2] + V.shape[2]]
        want = O.forward_project(H[:, :, :, :, sz], Vc)
        assert np.abs(img - want).max() < 1e-12 * np.abs(want).max()
    img0, _ = O.library_gpu_entry(recon, [0, 1], centre, 3, 1.5, H)
    img1, _ = O.library_gpu_entry(recon, [0, 1], centre, 3, 1.5, H, kernel=np.ones((1, 1, 1)))
    assert np.array_equal(img0, img1)
    kz = np.zeros((1, 1, 3))
    kz[0, 0, 0] = 1.0   # out[z] = V[z + 1]: the last plane of the sub-volume gets zero
    img2, p = O.library_gpu_entry(recon, [0, 1], centre, 3, 1.5, H, kernel=kz)
    sx, sy = slice(p["x0"] - 1, p["x1"]), slice(p["y0"] - 1, p["y1"])
    V = np.zeros((X, Y, p["z1"] - p["z0"] + 1))
    for l in range(V.shape[2] - 1):
        V[sx, sy, l] = 0.5 * (recon[0][sx, sy, p["z0"] + l] + recon[1][sx, sy, p["z0"] + l])
    want = O.forward_project(H[:, :, :, :, p["z0"] - 1:p["z1"]], V)
    assert np.abs(img2 - want).max() < 1e-12 * np.abs(want).max()


def test_matlab_round_half_away_from_zero():
    assert list(O.matlab_round([0.5, 1.5, 2.5, -0.5, -2.5, 2.4999])) == [1, 2, 3, -1, -3, 2]


# ---------------------------------------------------------------- library placement

def test_library_gpu_placement_hand_worked():
    p = O.library_gpu_placement((10.5, 3.2, 2.5), neur_rad=4, axial=2, image_size=(20, 20), nz=5)
    # round: x = 11, y = 3, z = 3; dz = ceil(4/2) = 2
    assert (p["x"], p["y"], p["z"]) == (11, 3, 3)
    assert (p["z0"], p["z1"]) == (1, 5)
    assert (p["x0"], p["x1"], p["y0"], p["y1"]) == (7, 15, 1, 7)
    p = O.library_gpu_placement((19.6, 19.4, 0.6), 3, 1.5, (20, 20), 4)
    assert (p["x"], p["y"], p["z"]) == (20, 19, 1)
    assert (p["z0"], p["z1"]) == (1, 3)
    assert (p["x0"], p["x1"], p["y0"], p["y1"]) == (17, 20, 16, 20)


def test_patched_rect_hand_worked():
    """project_forward_patched.m:8-24,54-70 worked by hand."""
    # pixels (10..12, 11..13), n 3, M 7: origin 9; n_ml = ceil(max/3) - 3 = (1, 2) -> (1, 3);
    # 'full' 9 x 15 centred at 9 + ceil(n_ml/2*3) = (11, 14) -> x 7..15, y 7..21, patch 1..9, 1..15
    g = O.patched_rect((10, 11, 1), (12, 13, 3), (30, 30), 3, 7)
    assert (g["x0"], g["x1"], g["y0"], g["y1"]) == (7, 15, 7, 21)
    assert (g["p0"], g["p1"], g["q0"], g["q1"]) == (1, 9, 1, 15)
    assert list(g["n_ml"]) == [1, 3] and (g["off_z"], g["n_z"]) == (0, 3)
    # at the sensor corner: centre 2, 'full' 9 wide -> x 1..6 (clipped), patch rows 4..9
    g = O.patched_rect((1, 2, 1), (2, 2, 1), (10, 10), 3, 7)
    assert (g["x0"], g["x1"], g["p0"], g["p1"]) == (1, 6, 4, 9)


def test_project_forward_patched_matches_full_sensor_projection():
    """The patched CPU library projection equals the full-sensor forward
    projection of the same voxels wherever the patch is pasted, for a patch
    well inside the sensor (project_forward_patched.m:54-73)."""
    n, M = 3, 7
    rng = np.random.default_rng(13)
    H = rng.random((M, M, n, n, 3))
    pixels = np.array([[10, 11, 2], [11, 11, 2], [12, 13, 3], [10, 12, 1]])
    colv = rng.random((2, 2, 2)) + 0.5
    idx = np.array([[1, 1, 1], [2, 1, 1], [2, 2, 2], [1, 2, 1]])
    img, info = O.project_forward_patched(pixels, (30, 30), n, H, colv, idx)
    V = np.zeros((30, 30, 3))
    for t in range(4):
        V[pixels[t, 0] - 1, pixels[t, 1] - 1, pixels[t, 2] - 1] = colv[idx[t, 0] - 1, idx[t, 1] - 1, idx[t, 2] - 1]
    full = np.zeros((30, 30))
    for c in range(3):
        Vc = np.zeros((30, 30, 3))
        Vc[:, :, c] = V[:, :, c]
        full += O.forward_project(H, Vc)
    # offsets: px_min = (10, 11, 1) -> off_ml = (3, 3), off_px = (9, 9); n_ml = ceil(max/3) - 3 = (1, 2) -> odd (1, 3)
    assert list(info["off_px"]) == [9, 9] and list(info["n_ml"]) == [1, 3]
    # the 'full' patch is pasted centred at off + ceil(n_ml/2 * n); compare on the pasted window
    x0, x1, y0, y1 = info["x0"], info["x1"], info["y0"], info["y1"]
    win = img[x0 - 1:x1, y0 - 1:y1]
    assert np.abs(win).sum() > 0
    # for odd n and M the paste centre off + ceil(n_ml*n/2) lands the 'full'
    # result exactly on the 'same'-aligned sensor position (zero shift)
    assert np.abs(full[x0 - 1:x1, y0 - 1:y1] - win).max() < 1e-12
    # outside the pasted rectangle the patched image is empty
    mask = np.ones_like(img, dtype=bool)
    mask[x0 - 1:x1, y0 - 1:y1] = False
    assert np.all(img[mask] == 0)


# ---------------------------------------------------------------- gradients

def test_poisson_nll_gradients_finite_differences():
    n, M = 3, 7
    rng = np.random.default_rng(14)
    H = rng.random((M, M, n, n, 2)) + 0.05
    V = rng.random((9, 9, 2)) + 0.2
    F = lambda v: O.forward_project(H, v)
    B = lambda o: O.backward_project(H, o)
    y = rng.poisson(F(V) * 3.0).astype(np.float64) / 3.0
    gV = O.poisson_nll_grad_volume(F, B, V, y)
    eps = 1e-6
    for (i, j, c) in [(0, 0, 0), (4, 5, 1), (8, 3, 0), (2, 7, 1)]:
        Vp, Vm = V.copy(), V.copy()
        Vp[i, j, c] += eps
        Vm[i, j, c] -= eps
        fd = (O.poisson_nll(F, Vp, y) - O.poisson_nll(F, Vm, y)) / (2 * eps)
        assert abs(fd - gV[i, j, c]) < 1e-6 * max(1.0, abs(fd))
    gH = O.poisson_nll_grad_psf(H, V, y)
    for (i, j, a, b, c) in [(3, 3, 0, 0, 0), (0, 6, 1, 2, 1), (5, 1, 2, 1, 0), (6, 6, 2, 2, 1)]:
        Hp, Hm = H.copy(), H.copy()
        Hp[i, j, a, b, c] += eps
        Hm[i, j, a, b, c] -= eps
        fd = (O.poisson_nll(lambda v: O.forward_project(Hp, v), V, y)
              - O.poisson_nll(lambda v: O.forward_project(Hm, v), V, y)) / (2 * eps)
        assert abs(fd - gH[i, j, a, b, c]) < 1e-6 * max(1.0, abs(fd))


def test_ls_deconv_decreases_objective():
    rng = np.random.default_rng(15)
    H = rng.random((7, 7, 3, 3, 2))
    V = rng.random((9, 12, 2))
    F = lambda v: O.forward_project(H, v)
    B = lambda o: O.backward_project(H, o)
    y = F(V)
    htf = B(y)
    objs = []
    for it in range(1, 6):
        X = O.ls_deconv(F, B, htf, max_iter=it)
        objs.append(np.linalg.norm(y - F(X)) ** 2)
        assert X.min() >= 0
    assert all(objs[i + 1] <= objs[i] * (1 + 1e-10) for i in range(len(objs) - 1))


def test_tv_l2_stencil_is_convn_same_of_the_reference_filter():
    """LS_deconv.m:54-58 builds a 3x3x3 filter and applies convn(X, filter, 'same');
    scipy's N-d 'same' convolution is the same operation for odd filters.
    Distinct multipliers per dimension catch a transposed axis."""
    from scipy.signal import convolve
    rng = np.random.default_rng(21)
    X = rng.random((7, 6, 5))
    lam = (0.3, -1.7, 2.9)
    filt = np.zeros((3, 3, 3))
    filt[1, 1, [0, 2]] = lam[2]   # filter(2,2,[1 3]) = lamb(3)
    filt[1, [0, 2], 1] = lam[1]   # filter(2,[1 3],2) = lamb(2)
    filt[[0, 2], 1, 1] = lam[0]   # filter([1 3],2,2) = lamb(1)
    filt[1, 1, 1] = -sum(lam)
    np.testing.assert_allclose(O.tv_l2_stencil(X, lam), convolve(X, filt, mode="same"), rtol=1e-12, atol=1e-12)


def _dense_problem(seed, rows=40, cols=15):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((rows, cols))
    x_true = np.maximum(rng.standard_normal(cols), 0)
    y = A @ x_true + 0.05 * rng.standard_normal(rows)
    return A, y


def test_ls_deconv_converges_to_nnls():
    """With every multiplier 0 the fixed point is the non-negative least-squares
    solution (scipy.optimize.nnls); with lamb_L2 it is NNLS of the stacked
    system [A; sqrt(lamb_L2) I] x = [y; 0]."""
    from scipy.optimize import nnls
    A, y = _dense_problem(5)
    F = lambda x: A @ x
    B = lambda o: A.T @ o
    X = O.ls_deconv(F, B, B(y), max_iter=3000)
    x_nnls, _ = nnls(A, y)
    assert np.abs(X - x_nnls).max() < 1e-7 * max(1.0, np.abs(x_nnls).max())
    l2 = 7.5
    X2 = O.ls_deconv(F, B, B(y), max_iter=3000, lamb_L2=l2)
    x2, _ = nnls(np.vstack([A, math.sqrt(l2) * np.eye(A.shape[1])]), np.concatenate([y, np.zeros(A.shape[1])]))
    assert np.abs(X2 - x2).max() < 1e-7 * max(1.0, np.abs(x2).max())
    assert np.abs(X2 - X).max() > 1e-3   # the L2 term changes the answer


def test_ls_deconv_l1_fixed_point_is_bound_constrained_minimum():
    """With lamb_L1 the iteration minimises 1/2||Ax-y||^2 + lamb_L1 sum(x) over
    x >= 0 (the gradient it follows is A'Ax - A'y + lamb_L1); compare with
    scipy's L-BFGS-B on that objective."""
    from scipy.optimize import minimize
    A, y = _dense_problem(6)
    lam = 2.0
    F = lambda x: A @ x
    B = lambda o: A.T @ o
    X = O.ls_deconv(F, B, B(y), max_iter=4000, lamb_L1=lam)
    f = lambda x: 0.5 * np.sum((A @ x - y) ** 2) + lam * np.sum(x)
    gr = lambda x: A.T @ (A @ x - y) + lam
    r = minimize(f, np.zeros(A.shape[1]), jac=gr, method="L-BFGS-B", bounds=[(0, None)] * A.shape[1],
                 options={"ftol": 1e-15, "gtol": 1e-12, "maxiter": 10000})
    assert abs(f(X) - r.fun) < 1e-9 * abs(r.fun)
    assert np.abs(X - r.x).max() < 1e-4
    assert (X == 0).sum() > 0   # the L1 term makes the solution sparse


@pytest.mark.parametrize("terms", [{}, {"lamb_L2": 0.7, "lamb_TV_L2": (0.2, 0.1, 0.3), "lamb_TV_L1": 1.0}])
def test_ls_deconv_step_is_exact_line_search(terms):
    """One iteration moves to X1 = max(df - a d, 0) where a minimises the
    quadratic model along -d.  With the default step the model is ||A x - y||^2
    only (LS_deconv.m:83-84); when lamb_TV_L1*lamb_TV_L2*lamb_L2 is nonzero in
    every entry it includes the penalties (:81).  A wrong denominator fails."""
    rng = np.random.default_rng(8)
    H = rng.random((7, 7, 3, 3, 2))
    F = lambda v: O.forward_project(H, v)
    B = lambda o: O.backward_project(H, o)
    y = F(rng.random((9, 12, 2)))
    df = B(y) * 1.3
    X1 = O.ls_deconv(F, B, df, max_iter=1, **terms)
    l2 = terms.get("lamb_L2", 0.0)
    tv = terms.get("lamb_TV_L2", (0.0, 0.0, 0.0))
    Qfull = lambda x: B(F(x)) + (O.tv_l2_stencil(x, tv) if max(tv) else 0) + l2 * x
    g = Qfull(df) - df
    d = np.where((df > 0) | (g < 0), g, 0.0)
    if terms:
        model = lambda x: 0.5 * np.sum(x * Qfull(x)) - np.sum((Qfull(df) - g) * x)
    else:
        # the default step's curvature along d is ||F(d)||^2 only
        model = lambda x: np.sum(g * (x - df)) + 0.5 * np.sum(F(x - df) ** 2)
    # the model is quadratic in a: fit it through three points and take its vertex
    s_ = 1e-3
    m0, m1, m2 = (model(df - k * s_ * d) for k in (0, 1, 2))
    c2 = (m2 - 2 * m1 + m0) / (2 * s_ * s_)
    c1 = (m1 - m0) / s_ - c2 * s_
    a = -c1 / (2 * c2)
    np.testing.assert_allclose(X1, np.maximum(df - a * d, 0), rtol=1e-6, atol=1e-9 * np.abs(df).max())


def test_convn_same_is_matlab_convn():
    """convn(V, k, 'same'): scipy's N-d 'same' convolution for odd kernels;
    for even sides MATLAB crops the full convolution from floor(size(k)/2)
    (its documented conv([1 2 3], [1 1], 'same') = [3 5 3])."""
    from scipy.signal import convolve
    rng = np.random.default_rng(31)
    V = rng.random((9, 8, 5))
    for ks in [(3, 3, 3), (5, 1, 3), (1, 1, 1), (3, 5, 7), (17, 17, 5)]:
        k = rng.random(ks)
        np.testing.assert_allclose(O.convn_same(V, k), convolve(V, k, mode="same"), rtol=1e-12, atol=1e-12)
    e = O.convn_same(np.array([1.0, 2.0, 3.0]).reshape(3, 1, 1), np.array([1.0, 1.0]).reshape(2, 1, 1))
    np.testing.assert_array_equal(e.ravel(), [3.0, 5.0, 3.0])
    k = rng.random((4, 2, 6))
    full = convolve(V, k, mode="full")
    o = [s // 2 for s in k.shape]
    np.testing.assert_allclose(O.convn_same(V, k), full[o[0]:o[0] + 9, o[1]:o[1] + 8, o[2]:o[2] + 5],
                               rtol=1e-12, atol=1e-12)
    # kernels more than twice the volume's side (taps shifted past the volume)
    W = rng.random((3, 2, 2))
    for ks in [(65, 1, 1), (7, 9, 5), (4, 6, 8)]:
        k = rng.random(ks)
        full = convolve(W, k, mode="full")
        o = [s // 2 for s in ks]
        np.testing.assert_allclose(O.convn_same(W, k), full[o[0]:o[0] + 3, o[1]:o[1] + 2, o[2]:o[2] + 2],
                                   rtol=1e-12, atol=1e-12)


def test_generate_kernel_known_answers():
    """src/generate_kernel.m: fspecial('gaussian', 3, 0.5) is MATLAB's documented
    [0.0113 0.0838 0.0113; 0.0838 0.6193 0.0838; ...]; a unit ball is the
    6-neighbour cross; every kernel has unit L2 norm; the Gaussian kernel is the
    lateral fspecial times exp(-z^2 / (4 p2^2)) on z = -3p2:3p2."""
    from sid_amd.recon import fspecial_gaussian, generate_kernel
    g = fspecial_gaussian(3, 0.5)
    np.testing.assert_allclose(g, [[0.0113, 0.0838, 0.0113], [0.0838, 0.6193, 0.0838], [0.0113, 0.0838, 0.0113]],
                               atol=5e-5)
    b = generate_kernel("ball", (1, 1))
    assert b.shape == (3, 3, 3)
    cross = np.zeros((3, 3, 3))
    cross[1, 1, :] = cross[1, :, 1] = cross[:, 1, 1] = 1
    np.testing.assert_allclose(b, cross / math.sqrt(7), rtol=1e-6)
    b2 = generate_kernel("ball", (8, 2))          # the find_kernel-sized neuron (17 x 17 x 5)
    assert b2.shape == (17, 17, 5) and abs(np.linalg.norm(b2) - 1) < 1e-6
    ga = generate_kernel("gaussian", (1.2, 1.0))
    assert ga.shape == (13, 13, 7)
    zf = np.exp(-np.arange(-3.0, 4.0) ** 2 / 4.0)
    ref = fspecial_gaussian(13, 1.2)[:, :, None] * zf[None, None, :]
    np.testing.assert_allclose(ga, ref / np.linalg.norm(ref), rtol=1e-6)
    lo = generate_kernel("lorentz", (0.5, 0.4))
    assert lo.shape == (7, 7, 5)
    c = (3, 3, 2)
    assert abs(lo[c] * np.linalg.norm(1.0 / (1 + 0.5 * (np.arange(-3, 4)[:, None, None] ** 2
                                                        + np.arange(-3, 4)[None, :, None] ** 2)
                                             + 0.4 * np.arange(-2, 3)[None, None, :] ** 2)) - 1.0) < 1e-6
    u = generate_kernel("user", np.ones((2, 3)))
    assert u.shape == (2, 3, 1) and abs(float(u.sum()) - math.sqrt(6)) < 1e-6


def test_ls_options_accept_the_config_default_scalar_tv():
    """sid_config_manage.m sets recon_opts.lamb_TV_L2 = 0 (a scalar); LS_deconv.m
    accepts it because max(0) leaves the TV term off, but a nonzero scalar fails
    at filter(2,2,[1 3]) = lamb_TV_L2(3)."""
    import sid_amd
    o = sid_amd._ls_opts({"lamb_TV_L2": 0})
    assert list(o.lamb_TV_L2) == [0.0, 0.0, 0.0] and o.max_iter == 8
    with pytest.raises(ValueError):
        sid_amd._ls_opts({"lamb_TV_L2": 0.3})


def test_rank_1_factorization_converges_to_the_leading_singular_pair():
    """src/rank_1_factorization.m is a power iteration on Y Y': after many
    iterations s t equals sigma_1 u_1 v_1' (numpy SVD); one iteration gives
    t = column sums / their norm."""
    rng = np.random.default_rng(41)
    Y = rng.random((60, 25)) + np.outer(rng.random(60), rng.random(25)) * 3
    s, t = O.rank_1_factorization(Y, 300)
    U, S, Vt = np.linalg.svd(Y, full_matrices=False)
    np.testing.assert_allclose(np.outer(s, t), S[0] * np.outer(U[:, 0], Vt[0]), rtol=1e-10, atol=1e-10)
    assert abs(np.linalg.norm(t) - 1) < 1e-12
    s1, t1 = O.rank_1_factorization(Y, 1)
    cs = Y.sum(axis=0)
    np.testing.assert_allclose(t1, cs / np.linalg.norm(cs), rtol=1e-12)


def test_compute_std_image_is_the_residual_std():
    """src/compute_std_image.m expands the per-pixel standard deviation (over
    frames, normalised by T) of Y - y1 y2 for the rank-1 factors (||y2|| = 1,
    y1 = Y y2'); with zero factors it is the plain std of Y."""
    rng = np.random.default_rng(42)
    Y = rng.random((50, 40)) * 2
    np.testing.assert_allclose(O.compute_std_image(Y), np.std(Y, axis=1), rtol=1e-10)
    s, t = O.rank_1_factorization(Y, 2)
    np.testing.assert_allclose(O.compute_std_image(Y, s, t), np.std(Y - np.outer(s, t), axis=1), rtol=1e-9)
