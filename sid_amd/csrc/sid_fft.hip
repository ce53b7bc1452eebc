// FFT path of the light-field projector for wide (out-of-focus) planes.
//
// The reference evaluates every (aa, bb, cc) term with a full N x N FFT
// convolution (external/LFrecon_util/conv2FFT_.m:8-15, N from
// forwardProjectGPU.m:9-13).  Here the pad size is chosen as N = n*N' with N'
// a 2/3/5-smooth length, N >= X + (kernel reach), so that the N-point DFT of a
// plane masked to the sublattice (a, b) is a phase times the N'-point DFT of
// the lenslet sub-image V_ab(Q) = V(a + nQ):
//     DFT_N(mask_ab V)(k) = e^{-2 pi i (kx a + ky b)/N} * DFT_N'(V_ab)(k mod N').
// With k = k0 + N' j (j = (jx, jy) in [0, n)^2) the forward projection is,
// for every base frequency k0, a complex GEMV over q = (plane, ab):
//     Ohat[k0][j] = sum_q G[k0][q][j] * Vt[k0][q],
//     G[k0][(c,ab)][j] = DFT_N(kernel H_abc placed at (i - off + a, j' - off + b) mod N)(k0 + N' j),
// followed by one inverse N x N transform.  The backward (adjoint) is
//     Bt[k0][q] = sum_j conj(G[k0][q][j]) * What(k0 + N' j),  B_ab = IDFT_N'(Bt) / n^2.
// Only k0y in [0, N'/2] is stored (real data: Hermitian symmetry gives the
// rest).  G is precomputed per handle and streamed from HBM once per
// projection in each direction: this path is HBM-bound.  In the BF16X2
// precision mode G is stored as FP16 pairs, row q scaled by 1/s_q with s_q the
// L1 norm of its kernel (the exact maximum of |G| for a non-negative kernel),
// which halves the stream; s_q is folded into the packed input (forward) and
// the unpacked output (backward).
//
// The FFTs are hand-written Stockham autosort transforms in LDS with
// radix-4/2/3/5 butterflies; twiddles come from tables computed on the host
// in double precision.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>

#include "sid_fft_core.h"
#include "sid_internal.h"

namespace sid {

// ---------------------------------------------------------------------------
// 1-D complex FFT, batched: block b transforms sequences b*spb .. b*spb+spb-1,
// each of length L (<= 2048), element t of sequence s at
// data[s * seq_stride + t * elem_stride].  dir = -1 forward, +1 inverse
// (unnormalised).  tw[t] = exp(dir * 2 pi i t / L) for t in [0, L).
// ---------------------------------------------------------------------------

template <int R>
__device__ __forceinline__ void stockham_stage(const float2 *__restrict__ in, float2 *__restrict__ out, int L, int Ns,
                                               const float2 *__restrict__ tw, float sgn, int spb, int tid, int nth)
{
    const int nb = L / R;
    const int step = L / (Ns * R);   // twiddle index step: angle = 2 pi k r / (Ns R); step * k * r < L
    const float rnb = 1.0f / (float)nb, rns = 1.0f / (float)Ns;
    for (int e = tid; e < spb * nb; e += nth) {
        const int s = qdiv(e, nb, rnb), j = e - s * nb;
        const float2 *src = in + s * L;
        float2 *dst = out + s * L;
        const int jq = qdiv(j, Ns, rns), k = j - jq * Ns;
        const int tstep = step * k;
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const float2 x = src[j + r * nb];
            v[r] = r == 0 ? x : cmul(x, tw[tstep * r]);
        }
        small_dft<R>(v, sgn);
        const int base = jq * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[base + r * Ns] = v[r];
    }
}

// Sequence s starts at (s / inner) * outer_stride + (s % inner) * seq_stride.
__device__ __forceinline__ int64_t seq_off(int64_t s, int64_t inner, int64_t outer_stride, int64_t seq_stride)
{
    return (s / inner) * outer_stride + (s % inner) * seq_stride;
}

__global__ void k_fft1d(float2 *__restrict__ data, FftDesc d, int64_t s_begin, int64_t nseq, int64_t inner,
                        int64_t outer_stride, int64_t seq_stride, int64_t elem_stride, const float2 *__restrict__ tw,
                        float sgn, float scale)
{
    extern __shared__ float2 sm[];
    __shared__ int64_t soff[64];   // start of each of the block's sequences (spb <= 64)
    float2 *b0 = sm, *b1 = sm + d.spb * d.L;
    const int64_t s0 = s_begin + (int64_t)blockIdx.x * d.spb;
    const int spb = (int)min((int64_t)d.spb, nseq - s0);
    if (spb <= 0) return;
    if ((int)threadIdx.x < spb) soff[threadIdx.x] = seq_off(s0 + threadIdx.x, inner, outer_stride, seq_stride);
    __syncthreads();
    const float rL = 1.0f / (float)d.L;
    for (int e = threadIdx.x; e < spb * d.L; e += blockDim.x) {
        const int s = qdiv(e, d.L, rL), t = e - s * d.L;
        b0[e] = data[soff[s] + (int64_t)t * elem_stride];
    }
    __syncthreads();
    int Ns = 1;
    float2 *in = b0, *out = b1;
    for (int st = 0; st < d.nst; ++st) {
        const int R = d.rad[st];
        if (R == 4) stockham_stage<4>(in, out, d.L, Ns, tw, sgn, spb, threadIdx.x, blockDim.x);
        else if (R == 2) stockham_stage<2>(in, out, d.L, Ns, tw, sgn, spb, threadIdx.x, blockDim.x);
        else if (R == 3) stockham_stage<3>(in, out, d.L, Ns, tw, sgn, spb, threadIdx.x, blockDim.x);
        else stockham_stage<5>(in, out, d.L, Ns, tw, sgn, spb, threadIdx.x, blockDim.x);
        __syncthreads();
        Ns *= R;
        float2 *t = in; in = out; out = t;
    }
    for (int e = threadIdx.x; e < spb * d.L; e += blockDim.x) {
        const int s = qdiv(e, d.L, rL), t = e - s * d.L;
        const float2 v = in[e];
        data[soff[s] + (int64_t)t * elem_stride] = make_float2(v.x * scale, v.y * scale);
    }
}

bool fft_factor(int L, std::vector<int> &rad)
{
    rad.clear();
    int m = L;
    while (m % 4 == 0) { rad.push_back(4); m /= 4; }
    while (m % 2 == 0) { rad.push_back(2); m /= 2; }
    while (m % 3 == 0) { rad.push_back(3); m /= 3; }
    while (m % 5 == 0) { rad.push_back(5); m /= 5; }
    return m == 1 && rad.size() <= 12;
}

bool is_smooth235(int L)
{
    std::vector<int> r;
    return L >= 1 && fft_factor(L, r);
}

int next_smooth235(int L)
{
    while (!is_smooth235(L)) ++L;
    return L;
}

int fft_plan_make(FftPlan1D &pl, int L)
{
    std::vector<int> rad;
    if (L < 1 || L > 2048 || !fft_factor(L, rad)) {
        set_error("FFT length %d is not a 2/3/5-smooth length <= 2048", L);
        return SID_ERR_DIMS;
    }
    pl.L = L;
    pl.nst = (int)rad.size();
    for (int i = 0; i < pl.nst; ++i) pl.rad[i] = rad[i];
    std::vector<float2> tf(L), ti(L);
    for (int t = 0; t < L; ++t) {
        const double a = 2.0 * M_PI * (double)t / (double)L;
        tf[t] = make_float2((float)std::cos(a), (float)-std::sin(a));
        ti[t] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    SID_TRY_HIP(hipMalloc(&pl.tw_fwd, sizeof(float2) * L));
    SID_TRY_HIP(hipMalloc(&pl.tw_inv, sizeof(float2) * L));
    SID_TRY_HIP(hipMemcpy(pl.tw_fwd, tf.data(), sizeof(float2) * L, hipMemcpyHostToDevice));
    SID_TRY_HIP(hipMemcpy(pl.tw_inv, ti.data(), sizeof(float2) * L, hipMemcpyHostToDevice));
    return SID_OK;
}

void fft_plan_free(FftPlan1D &pl)
{
    if (pl.tw_fwd) (void)hipFree(pl.tw_fwd);
    if (pl.tw_inv) (void)hipFree(pl.tw_inv);
    pl.tw_fwd = pl.tw_inv = nullptr;
    pl.L = 0;
}

static FftDesc fft_desc(const FftPlan1D &pl)
{
    FftDesc d;
    d.L = pl.L;
    d.nst = pl.nst;
    for (int i = 0; i < 12; ++i) d.rad[i] = i < pl.nst ? pl.rad[i] : 1;
    d.spb = std::max(1, std::min(64, 1024 / pl.L));
    return d;
}

// Batched 1-D transforms; dir -1 forward, +1 inverse (scaled by `scale`).
// Sequence s (0 <= s < nseq) starts at (s / inner) * outer_stride + (s % inner) * seq_stride.
int fft_run(const FftPlan1D &pl, float2 *data, int64_t nseq, int64_t inner, int64_t outer_stride,
            int64_t seq_stride, int64_t elem_stride, int dir, float scale, hipStream_t st)
{
    if (nseq <= 0) return SID_OK;
    FftDesc d = fft_desc(pl);
    const int64_t nblk = (nseq + d.spb - 1) / d.spb;
    const size_t lds = sizeof(float2) * 2 * d.spb * pl.L;
    for (int64_t b0 = 0; b0 < nblk; b0 += 65535) {
        const int64_t nb = std::min<int64_t>(65535, nblk - b0);
        hipLaunchKernelGGL(k_fft1d, dim3((unsigned)nb), dim3(256), lds, st, data, d, b0 * d.spb, nseq,
                           std::max<int64_t>(inner, 1), outer_stride, seq_stride, elem_stride,
                           dir < 0 ? pl.tw_fwd : pl.tw_inv, dir < 0 ? -1.0f : 1.0f, scale);
        SID_TRY_HIP(hipGetLastError());
    }
    return SID_OK;
}

// ---------------------------------------------------------------------------
// FFT projector kernels
// ---------------------------------------------------------------------------

// Lenslet sub-images of the FFT planes: subs[q][Qx][Qy], q = slot * n2 + ab.
__global__ void k_fft_sub_gather(const float *__restrict__ vol, const int *__restrict__ slot_plane, int nq, int n2,
                                 int Nxp, int Nyp, int Lx, int Ly, int64_t XP, int64_t YP, int T, int S, int nz_in,
                                 float2 *__restrict__ subs)
{
    const int64_t total = (int64_t)nq * Nxp * Nyp;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int Qy = (int)(e % Nyp);
        const int64_t t = e / Nyp;
        const int Qx = (int)(t % Nxp), q = (int)(t / Nxp);
        const int c = slot_plane[q / n2], ab = q % n2;
        float v = 0.0f;
        if (Qx < Lx && Qy < Ly && c < nz_in) v = vol[(((int64_t)c * XP + Qx + T) * YP + (Qy + T)) * S + ab];
        subs[e] = make_float2(v, 0.0f);
    }
}

// vt[k0][q] = s_q * subs[q][k0x][k0y], k0 = k0x * Hy + k0y (s_q = 1 without qscale).
__global__ void k_fft_pack_vt(const float2 *__restrict__ subs, const float *__restrict__ qscale, int nq, int Nxp,
                              int Nyp, int Hy, float2 *__restrict__ vt)
{
    const int64_t total = (int64_t)Nxp * Hy * nq;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int q = (int)(e % nq);
        const int k0 = (int)(e / nq);
        const int k0x = k0 / Hy, k0y = k0 - k0x * Hy;
        float2 v = subs[((int64_t)q * Nxp + k0x) * Nyp + k0y];
        if (qscale) {
            const float sq = qscale[q];
            v.x *= sq;
            v.y *= sq;
        }
        vt[e] = v;
    }
}

__device__ __forceinline__ float2 spec_ld(const float2 *p) { return *p; }
__device__ __forceinline__ float2 spec_ld(const __half2 *p) { return __half22float2(*p); }

// Loads of one tile of nel elements at src (16-byte aligned) into registers:
// vector u of this thread is element range [(u * 256 + tid) * EPV, +EPV);
// vectors past the tile re-read its first vector (always inside the spectra
// buffer, whose k0 blocks hold at least 16 bytes) and are not stored.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <typename GT, int kVec>
__device__ __forceinline__ void tile_fetch(u32x4 (&pre)[kVec], GT &tail, const GT *__restrict__ src, int nel)
{
    constexpr int EPV = 16 / (int)sizeof(GT);
    const int nv = nel / EPV;
    const u32x4 *s4 = reinterpret_cast<const u32x4 *>(src);
#pragma unroll
    for (int u = 0; u < kVec; ++u) {
        const int e = u * 256 + (int)threadIdx.x;
        // the spectra are streamed once per projection (tens of GB, far beyond
        // the last-level cache): non-temporal loads measure ~10 % faster
        pre[u] = __builtin_nontemporal_load(s4 + (e < nv ? e : 0));
    }
    const int et = nv * EPV + (int)threadIdx.x;
    if (et < nel) tail = src[et];
}

// Forward GEMV: oh[s][k0][j] = sum_{q in chunk s} G[k0][q][j] * vt[k0][q];
// block (k0, s), thread j.  The q range of a k0 is split over gridDim.y
// blocks so the grid is many times the resident capacity (no half-empty last
// round); k_fft_oh_sum adds the chunks in a fixed order.  The chunk is one
// contiguous stretch of HBM, streamed in tiles of 16 rows with 16-byte loads
// through registers (the next tile's loads are in flight while the current
// one is summed from LDS); reading rows directly (64 consecutive elements per
// wave) fetches every cache line that straddles two waves' pieces twice.
// vt[k0][q] is the same for every lane (scalar loads).
template <typename GT>
__global__ __launch_bounds__(256) void k_fft_gemv_fwd(const GT *__restrict__ G, const float2 *__restrict__ vt,
                                                      int nq, int n2, int qc, int nk0, int64_t kstride,
                                                      float2 *__restrict__ oh)
{
    constexpr int TQ = 16;
    constexpr int EPV = 16 / (int)sizeof(GT);
    constexpr int kVec = TQ * 256 / EPV / 256;
    __shared__ u32x4 tile[TQ * 256 / EPV];
    const int k0 = blockIdx.x;
    const int j = threadIdx.x;
    const int qb = blockIdx.y * qc, qe = min(nq, qb + qc);
    const GT *gk = G + (int64_t)k0 * kstride;
    const float2 *v = vt + (int64_t)k0 * nq;
    const GT *ts = reinterpret_cast<const GT *>(tile);
    u32x4 pre[kVec];
    GT pre_tail = GT();
    float2 acc0 = make_float2(0.f, 0.f), acc1 = make_float2(0.f, 0.f);
    if (qb < qe) tile_fetch<GT, kVec>(pre, pre_tail, gk + (int64_t)qb * n2, min(TQ, qe - qb) * n2);
    for (int q0 = qb; q0 < qe; q0 += TQ) {
        const int nr = min(TQ, qe - q0);
        __syncthreads();   // previous tile consumed
        {
            const int nel = nr * n2, nv = nel / EPV;
#pragma unroll
            for (int u = 0; u < kVec; ++u) {
                const int e = u * 256 + (int)threadIdx.x;
                if (e < nv) tile[e] = pre[u];
            }
            const int et = nv * EPV + (int)threadIdx.x;
            if (et < nel) reinterpret_cast<GT *>(tile)[et] = pre_tail;
        }
        __syncthreads();
        if (q0 + TQ < qe)
            tile_fetch<GT, kVec>(pre, pre_tail, gk + (int64_t)(q0 + TQ) * n2, min(TQ, qe - q0 - TQ) * n2);
        if (j < n2) {
            if (nr == TQ) {
#pragma unroll
                for (int r = 0; r < TQ; r += 2) {
                    const float2 g0 = spec_ld(ts + r * n2 + j), v0 = v[q0 + r];
                    const float2 g1 = spec_ld(ts + (r + 1) * n2 + j), v1 = v[q0 + r + 1];
                    acc0.x += g0.x * v0.x - g0.y * v0.y;
                    acc0.y += g0.x * v0.y + g0.y * v0.x;
                    acc1.x += g1.x * v1.x - g1.y * v1.y;
                    acc1.y += g1.x * v1.y + g1.y * v1.x;
                }
            } else {
                for (int r = 0; r < nr; ++r) {
                    const float2 g0 = spec_ld(ts + r * n2 + j), v0 = v[q0 + r];
                    acc0.x += g0.x * v0.x - g0.y * v0.y;
                    acc0.y += g0.x * v0.y + g0.y * v0.x;
                }
            }
        }
    }
    if (j < n2) oh[((int64_t)blockIdx.y * nk0 + k0) * n2 + j] = make_float2(acc0.x + acc1.x, acc0.y + acc1.y);
}

// Chunk sums of the forward GEMV, transposed: ohT[j][k0] = sum_s oh[s][k0][j].
// 32 x 32 tiles through LDS so reads (along j) and writes (along k0) are both
// contiguous; chunks are added in a fixed order.
__global__ __launch_bounds__(256) void k_fft_oh_sum(const float2 *__restrict__ oh, int nsplit, int nk0, int n2,
                                                    float2 *__restrict__ ohT)
{
    __shared__ float2 t[32][33];
    const int k0b = blockIdx.x * 32, jb = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8 threads
    for (int r = ty; r < 32; r += 8) {
        const int k0 = k0b + r, j = jb + tx;
        float2 v = make_float2(0.f, 0.f);
        if (k0 < nk0 && j < n2)
            for (int sp = 0; sp < nsplit; ++sp) {
                const float2 u = oh[((int64_t)sp * nk0 + k0) * n2 + j];
                v.x += u.x;
                v.y += u.y;
            }
        t[r][tx] = v;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int j = jb + r, k0 = k0b + tx;
        if (k0 < nk0 && j < n2) ohT[(int64_t)j * nk0 + k0] = t[tx][r];
    }
}

// Full N x N spectrum of the forward output from ohT[j][k0] (stored for
// k0y < Hy) by Hermitian symmetry; j = jx + n * jy.  Consecutive ky read
// consecutive k0.
__global__ void k_fft_assemble(const float2 *__restrict__ ohT, int nk0, int Nx, int Ny, int Nxp, int Nyp, int Hy,
                               int n, float2 *__restrict__ grid)
{
    const int64_t total = (int64_t)Nx * Ny;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int kx = (int)(e / Ny), ky = (int)(e % Ny);
        int ax = kx, ay = ky;
        bool cj = false;
        if (ky % Nyp >= Hy) {
            ax = (Nx - kx) % Nx;
            ay = (Ny - ky) % Ny;
            cj = true;
        }
        const int k0 = (ax % Nxp) * Hy + (ay % Nyp);
        const int j = ax / Nxp + n * (ay / Nyp);
        float2 v = ohT[(int64_t)j * nk0 + k0];
        if (cj) v.y = -v.y;
        grid[e] = v;
    }
}

// Sensor image (padded lenslet layout) -> complex N x N grid (zero padded).
__global__ void k_fft_img_grid(const float *__restrict__ img, int Nx, int Ny, int Xn, int Yn, int n, int64_t YP,
                               int T, int S, float2 *__restrict__ grid)
{
    const int64_t total = (int64_t)Nx * Ny;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(e / Ny), y = (int)(e % Ny);
        float v = 0.0f;
        if (x < Xn && y < Yn)
            v = img[((int64_t)(x / n + T) * YP + (y / n + T)) * S + (x % n) + n * (y % n)];
        grid[e] = make_float2(v, 0.0f);
    }
}

// Lane l of each 16-lane row gets lane l - o's value (0 for l < o), o in {1, 2, 4, 8}.
__device__ __forceinline__ float row_shr_f(float v, int o)
{
    const int x = __float_as_int(v);
    int y;
    switch (o) {
        case 8: y = __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true); break;
        case 4: y = __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true); break;
        case 2: y = __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true); break;
        default: y = __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true); break;
    }
    return __int_as_float(y);
}

// Backward GEMV: bt[k0][q] = sum_j conj(G[k0][q][j]) * W(k0 + N' j).  Block
// (k0, s) takes rows q of chunk s (qc rows, a multiple of 16).  The chunk is
// contiguous in HBM, so tiles of 16 rows are copied into LDS with 16-byte
// loads (every tile starts 16-byte aligned: k0 blocks are padded to a multiple
// of 16 bytes, a tile starts at a multiple of 16 rows, and the host checks the
// base alignment); then 16 lanes sum
// one row (lane l takes j = l + 16 t, conflict-free LDS reads, its 16 image
// values W held in registers) and reduce with four DPP row shifts.  Results
// for consecutive q are stored contiguously (k_fft_unpack_bt moves them into
// the per-q sub-spectra).  Requires n*n <= 256.
template <typename GT>
__global__ __launch_bounds__(256) void k_fft_gemv_bwd(const GT *__restrict__ G, const float2 *__restrict__ grid,
                                                      int nq, int qc, int64_t kstride, int n, int Nx, int Ny, int Nxp,
                                                      int Nyp, int Hy, float2 *__restrict__ bt)
{
    constexpr int TQ = 16;                      // rows per tile
    constexpr int EPV = 16 / (int)sizeof(GT);   // elements per 16-byte load
    __shared__ u32x4 tile[TQ * 256 / EPV];
    __shared__ float2 w[256];
    const int n2 = n * n;
    const int k0 = blockIdx.x;
    const int k0x = k0 / Hy, k0y = k0 - k0x * Hy;
    {
        const int j = threadIdx.x;
        float2 v = make_float2(0.f, 0.f);
        if (j < n2) {
            const int jx = j % n, jy = j / n;
            v = grid[(int64_t)(k0x + Nxp * jx) * Ny + (k0y + Nyp * jy)];
        }
        w[j] = v;
    }
    const int qb = blockIdx.y * qc, qe = min(nq, qb + qc);
    const GT *gk = G + (int64_t)k0 * kstride;
    const GT *ts = reinterpret_cast<const GT *>(tile);
    const int r = threadIdx.x >> 4, sl = threadIdx.x & 15;
    // lane sl's image values W(k0 + N' j), j = sl + 16 t, held in registers for
    // every row of the block (only the spectra tile is read from LDS per row)
    __syncthreads();
    float2 wr[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) wr[t] = w[sl + 16 * t];
    // Tiles are staged through registers: the loads of tile t+1 are issued
    // before tile t is summed, so HBM reads overlap the LDS work.  A tile is
    // at most 16 * 256 elements = kVec 16-byte vectors per thread.
    constexpr int kVec = TQ * 256 / EPV / 256;
    u32x4 pre[kVec];
    GT pre_tail = GT();
    if (qb < qe) tile_fetch<GT, kVec>(pre, pre_tail, gk + (int64_t)qb * n2, min(TQ, qe - qb) * n2);
    for (int q0 = qb; q0 < qe; q0 += TQ) {
        const int nr = min(TQ, qe - q0);
        __syncthreads();   // previous tile consumed (and w written, first pass)
        {
            const int nel = nr * n2, nv = nel / EPV;
#pragma unroll
            for (int u = 0; u < kVec; ++u) {
                const int e = u * 256 + (int)threadIdx.x;
                if (e < nv) tile[e] = pre[u];
            }
            const int et = nv * EPV + (int)threadIdx.x;
            if (et < nel) reinterpret_cast<GT *>(tile)[et] = pre_tail;
        }
        __syncthreads();
        if (q0 + TQ < qe)
            tile_fetch<GT, kVec>(pre, pre_tail, gk + (int64_t)(q0 + TQ) * n2, min(TQ, qe - q0 - TQ) * n2);
        if (r < nr) {
            float re = 0.f, im = 0.f;
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const int j = sl + 16 * t;
                if (j < n2) {
                    const float2 a = spec_ld(ts + r * n2 + j), b = wr[t];
                    re += a.x * b.x + a.y * b.y;   // conj(a) * b
                    im += a.x * b.y - a.y * b.x;
                }
            }
            // sum over the row's 16 lanes with DPP row shifts (no LDS traffic):
            // lane 15 of the row ends with the total
#pragma unroll
            for (int o = 8; o > 0; o >>= 1) {
                re += row_shr_f(re, o);
                im += row_shr_f(im, o);
            }
            if (sl == 15) bt[(int64_t)k0 * nq + q0 + r] = make_float2(re, im);
        }
    }
}

// s_q * bt[k0][q] -> subs[q][k0x][k0y] (k0y < Hy; s_q = 1 without qscale).
__global__ void k_fft_unpack_bt(const float2 *__restrict__ bt, const float *__restrict__ qscale, int nq, int Nxp,
                                int Nyp, int Hy, float2 *__restrict__ subs)
{
    const int64_t total = (int64_t)Nxp * Hy * nq;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int q = (int)(e / ((int64_t)Nxp * Hy));
        const int k0 = (int)(e - (int64_t)q * Nxp * Hy);
        const int k0x = k0 / Hy, k0y = k0 - k0x * Hy;
        float2 v = bt[(int64_t)k0 * nq + q];
        if (qscale) {
            const float sq = qscale[q];
            v.x *= sq;
            v.y *= sq;
        }
        subs[((int64_t)q * Nxp + k0x) * Nyp + k0y] = v;
    }
}

// Fill ky in [Hy, Nyp) of every sub-spectrum by Hermitian symmetry.
__global__ void k_fft_herm_expand(float2 *__restrict__ subs, int nq, int Nxp, int Nyp, int Hy)
{
    const int nhi = Nyp - Hy;
    const int64_t total = (int64_t)nq * Nxp * nhi;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int ky = Hy + (int)(e % nhi);
        const int64_t t = e / nhi;
        const int kx = (int)(t % Nxp), q = (int)(t / Nxp);
        const float2 v = subs[((int64_t)q * Nxp + (Nxp - kx) % Nxp) * Nyp + (Nyp - ky)];
        subs[((int64_t)q * Nxp + kx) * Nyp + ky] = make_float2(v.x, -v.y);
    }
}

// Real parts of the inverse sub-transforms -> backward result in the padded
// volume layout (FFT planes only).
__global__ void k_fft_sub_scatter(const float2 *__restrict__ subs, const int *__restrict__ slot_plane, int nq, int n2,
                                  int Nxp, int Nyp, int Lx, int Ly, int64_t XP, int64_t YP, int T, int S,
                                  float *__restrict__ dfft)
{
    const int64_t total = (int64_t)nq * Lx * Ly;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int Qy = (int)(e % Ly);
        const int64_t t = e / Ly;
        const int Qx = (int)(t % Lx), q = (int)(t / Lx);
        const int c = slot_plane[q / n2], ab = q % n2;
        dfft[(((int64_t)c * XP + Qx + T) * YP + (Qy + T)) * S + ab] = subs[((int64_t)q * Nxp + Qx) * Nyp + Qy].x;
    }
}

// ---------------------------------------------------------------------------
// Fused sub-image transforms.  One block per pair of sub-images (q, q+1)
// transformed together as z = a + i b (two real 2-D DFTs for the price of one
// complex one); the Nx' x Ny' grid and its ping-pong copy live in LDS.
// ---------------------------------------------------------------------------

struct SubGeom {
    int nq, n2, Nxp, Nyp, Hy, Lx, Ly, T, S, nz_in;
    int64_t XP, YP;
};

// Row pitch of a sub-image grid in LDS: Ny' rounded up to odd, so that lanes
// reading neighbouring rows at one column hit different banks.
__host__ __device__ __forceinline__ int sub_pitch(int Nyp) { return Nyp | 1; }

// Forward: volume -> 2-D DFT of the pair's sub-images -> vt[k0][q] = s_q * V_q(k0)
// for the stored half spectrum.  Z = DFT(a + i b) gives
//   A(k) = (Z(k) + conj(Z(-k))) / 2,   B(k) = (Z(k) - conj(Z(-k))) / (2 i).
constexpr int kSubThreads = 1024;   // one block per CU (the two grids take most of the LDS)
constexpr int kSubUnroll = 8;       // grid elements per thread (prefetch registers): grids up to 8192 points

// Pairs of one block.  The kernels are persistent (one block per CU); blocks
// are dealt to the 8 XCDs round robin and each XCD has its own L2.
// Neighbouring pairs read the same cache lines (the volume's lenslet rows, the
// packed spectra), so each XCD takes one contiguous range of pairs, and that
// range is shared by the XCD's blocks in an interleaved order.
struct PairIter {
    int next, end, step;
};

__device__ __forceinline__ PairIter pair_iter(int npairs)
{
    constexpr int kXcd = 8;
    const int x = blockIdx.x % kXcd, i = blockIdx.x / kXcd;
    const int q = npairs / kXcd, r = npairs % kXcd;
    const int begin = x * q + min(x, r), count = q + (x < r ? 1 : 0);
    const int nbx = ((int)gridDim.x - x + kXcd - 1) / kXcd;   // blocks on this XCD
    return PairIter{begin + i, begin + count, nbx};
}

// Twiddle tables of both lengths into LDS (after the two grids).
__device__ __forceinline__ void load_twiddles(float2 *tx, float2 *ty, const float2 *__restrict__ twx,
                                              const float2 *__restrict__ twy, int Lx, int Ly)
{
    for (int t = threadIdx.x; t < Lx; t += blockDim.x) tx[t] = twx[t];
    for (int t = threadIdx.x; t < Ly; t += blockDim.x) ty[t] = twy[t];
}

// Forward: volume -> 2-D DFT of the pair's sub-images -> vt[k0][q] = s_q * V_q(k0)
// for the stored half spectrum.  Z = DFT(a + i b) gives
//   A(k) = (Z(k) + conj(Z(-k))) / 2,   B(k) = (Z(k) - conj(Z(-k))) / (2 i).
// The next pair's volume values are loaded into registers while the current
// pair is transformed and written out.
// ... and two adjacent complex values (q0, q0 + 1) of vt as one 16-byte access
__device__ __forceinline__ void ld_c2(const float2 *p, float2 &x, float2 &y)
{
    float4 v;
    __builtin_memcpy(&v, p, 16);
    x = make_float2(v.x, v.y);
    y = make_float2(v.z, v.w);
}
__device__ __forceinline__ void st_c2(float2 *p, float2 x, float2 y)
{
    const float4 v = make_float4(x.x, x.y, y.x, y.y);
    __builtin_memcpy(p, &v, 16);
}

__device__ __forceinline__ void sub_fwd_fetch(const float *__restrict__ vol, const int *__restrict__ slot_plane,
                                              const SubGeom &g, int pair, float (&a)[kSubUnroll],
                                              float (&b)[kSubUnroll])
{
    const int q0 = 2 * pair;
    const bool has1 = q0 + 1 < g.nq;
    const int c0 = slot_plane[q0 / g.n2], ab0 = q0 % g.n2;
    const int c1 = has1 ? slot_plane[(q0 + 1) / g.n2] : 0, ab1 = has1 ? (q0 + 1) % g.n2 : 0;
    const bool in0 = c0 < g.nz_in, in1 = has1 && c1 < g.nz_in;
    const float *v0p = vol + (int64_t)c0 * g.XP * g.YP * g.S + ab0;
    const float *v1p = vol + (int64_t)c1 * g.XP * g.YP * g.S + ab1;
    // the two sub-images are neighbours in one lenslet record: one load for both
    const bool adjacent = in0 && in1 && c1 == c0;
    const int npix = g.Nxp * g.Nyp;
    const float rnyp = 1.0f / (float)g.Nyp;
    // adjacency is uniform per pair: separate loops keep the 8-byte loads whole
    if (adjacent) {
#pragma unroll
        for (int u = 0; u < kSubUnroll; ++u) {
            const int e = u * kSubThreads + (int)threadIdx.x;
            const int Qx = qdiv(e, g.Nyp, rnyp), Qy = e - Qx * g.Nyp;
            float2 v = make_float2(0.f, 0.f);
            if (e < npix && Qx < g.Lx && Qy < g.Ly)
                v = ld_f2_a4(v0p + (((int64_t)Qx + g.T) * g.YP + (Qy + g.T)) * g.S);
            a[u] = v.x;
            b[u] = v.y;
        }
        return;
    }
#pragma unroll
    for (int u = 0; u < kSubUnroll; ++u) {
        const int e = u * kSubThreads + (int)threadIdx.x;
        const int Qx = qdiv(e, g.Nyp, rnyp), Qy = e - Qx * g.Nyp;
        a[u] = b[u] = 0.f;
        if (e < npix && Qx < g.Lx && Qy < g.Ly) {
            const int64_t cell = (((int64_t)Qx + g.T) * g.YP + (Qy + g.T)) * g.S;
            if (in0) a[u] = v0p[cell];
            if (in1) b[u] = v1p[cell];
        }
    }
}

__global__ __launch_bounds__(kSubThreads) void k_fft_sub_fwd(const float *__restrict__ vol,
                                                             const int *__restrict__ slot_plane,
                                                             const float *__restrict__ qscale, SubGeom g, FftDesc dx,
                                                             FftDesc dy, const float2 *__restrict__ twx,
                                                             const float2 *__restrict__ twy, float2 *__restrict__ vt)
{
    extern __shared__ float2 sm[];
    const int npix = g.Nxp * g.Nyp, P = sub_pitch(g.Nyp), nlds = g.Nxp * P;
    float2 *bufa = sm, *bufb = sm + nlds, *tx = sm + 2 * nlds, *ty = tx + g.Nxp;
    load_twiddles(tx, ty, twx, twy, g.Nxp, g.Nyp);
    PairIter it = pair_iter((g.nq + 1) / 2);
    if (it.next >= it.end) return;   // block-uniform
    const float rhy = 1.0f / (float)g.Hy, rnyp = 1.0f / (float)g.Nyp;
    const int nk0 = g.Nxp * g.Hy;
    float a[kSubUnroll], b[kSubUnroll];
    sub_fwd_fetch(vol, slot_plane, g, it.next, a, b);
    for (int pair = it.next; pair < it.end; pair += it.step) {
        lds_barrier();   // the previous pair's grids are consumed (twiddles loaded, first pass)
#pragma unroll
        for (int u = 0; u < kSubUnroll; ++u) {
            const int e = u * kSubThreads + (int)threadIdx.x;
            if (e < npix) {
                const int Qx = qdiv(e, g.Nyp, rnyp), Qy = e - Qx * g.Nyp;
                bufa[Qx * P + Qy] = make_float2(a[u], b[u]);
            }
        }
        lds_barrier();
        if (pair + it.step < it.end) sub_fwd_fetch(vol, slot_plane, g, pair + it.step, a, b);
        float2 *r1 = fft_pass_2d(bufa, bufb, dy, ty, -1.0f, g.Nxp, P, 1);               // along y (rows)
        float2 *r2 = fft_pass_2d(r1, r1 == bufa ? bufb : bufa, dx, tx, -1.0f, g.Nyp, 1, P);   // along x
        const int q0 = 2 * pair;
        const bool has1 = q0 + 1 < g.nq;
        const float s0 = qscale ? qscale[q0] : 1.0f;
        const float s1 = (qscale && has1) ? qscale[q0 + 1] : 1.0f;
        if (has1) {   // uniform per pair: one 16-byte store of both spectra per k0
            for (int k0 = threadIdx.x; k0 < nk0; k0 += blockDim.x) {
                const int kx = qdiv(k0, g.Hy, rhy), ky = k0 - kx * g.Hy;
                const float2 z = r2[kx * P + ky];
                const int mx = kx == 0 ? 0 : g.Nxp - kx, my = ky == 0 ? 0 : g.Nyp - ky;
                const float2 zm = r2[mx * P + my];
                const float2 av = make_float2(0.5f * (z.x + zm.x), 0.5f * (z.y - zm.y));
                const float2 bv = make_float2(0.5f * (z.y + zm.y), -0.5f * (z.x - zm.x));
                st_c2(vt + (int64_t)k0 * g.nq + q0, make_float2(av.x * s0, av.y * s0),
                      make_float2(bv.x * s1, bv.y * s1));
            }
        } else {
            for (int k0 = threadIdx.x; k0 < nk0; k0 += blockDim.x) {
                const int kx = qdiv(k0, g.Hy, rhy), ky = k0 - kx * g.Hy;
                const float2 z = r2[kx * P + ky];
                const int mx = kx == 0 ? 0 : g.Nxp - kx, my = ky == 0 ? 0 : g.Nyp - ky;
                const float2 zm = r2[mx * P + my];
                vt[(int64_t)k0 * g.nq + q0] = make_float2(0.5f * (z.x + zm.x) * s0, 0.5f * (z.y - zm.y) * s0);
            }
        }
    }
}

// Backward: s_q * bt[k0][q] (stored half spectrum, Hermitian completion) ->
// inverse 2-D DFT of A + i B -> real parts a, b into the padded volume layout,
// scaled by `scale`.  The next pair's spectra are loaded into registers while
// the current pair is transformed and written out.
__device__ __forceinline__ void sub_bwd_fetch(const float2 *__restrict__ bt, const SubGeom &g, int pair,
                                              float2 (&a)[kSubUnroll], float2 (&b)[kSubUnroll],
                                              unsigned &cjmask)
{
    const int q0 = 2 * pair;
    const bool has1 = q0 + 1 < g.nq;
    const int npix = g.Nxp * g.Nyp;
    const float rnyp = 1.0f / (float)g.Nyp;
    cjmask = 0u;
    // has1 is uniform per pair: separate loops keep the 16-byte loads whole
    if (has1) {
#pragma unroll
        for (int u = 0; u < kSubUnroll; ++u) {
            const int e = u * kSubThreads + (int)threadIdx.x;
            a[u] = b[u] = make_float2(0.f, 0.f);
            if (e < npix) {
                const int kx = qdiv(e, g.Nyp, rnyp), ky = e - kx * g.Nyp;
                int sx = kx, sy = ky;
                if (ky >= g.Hy) {   // conj of the stored value at -k
                    sx = kx == 0 ? 0 : g.Nxp - kx;
                    sy = g.Nyp - ky;
                    cjmask |= 1u << u;
                }
                ld_c2(bt + (int64_t)(sx * g.Hy + sy) * g.nq + q0, a[u], b[u]);
            }
        }
        return;
    }
#pragma unroll
    for (int u = 0; u < kSubUnroll; ++u) {
        const int e = u * kSubThreads + (int)threadIdx.x;
        a[u] = b[u] = make_float2(0.f, 0.f);
        if (e < npix) {
            const int kx = qdiv(e, g.Nyp, rnyp), ky = e - kx * g.Nyp;
            int sx = kx, sy = ky;
            if (ky >= g.Hy) {
                sx = kx == 0 ? 0 : g.Nxp - kx;
                sy = g.Nyp - ky;
                cjmask |= 1u << u;
            }
            a[u] = bt[(int64_t)(sx * g.Hy + sy) * g.nq + q0];
        }
    }
}

__global__ __launch_bounds__(kSubThreads) void k_fft_sub_bwd(const float2 *__restrict__ bt,
                                                             const int *__restrict__ slot_plane,
                                                             const float *__restrict__ qscale, SubGeom g, FftDesc dx,
                                                             FftDesc dy, const float2 *__restrict__ twx,
                                                             const float2 *__restrict__ twy, float scale,
                                                             float *__restrict__ dfft)
{
    extern __shared__ float2 sm[];
    const int npix = g.Nxp * g.Nyp, P = sub_pitch(g.Nyp), nlds = g.Nxp * P;
    float2 *bufa = sm, *bufb = sm + nlds, *tx = sm + 2 * nlds, *ty = tx + g.Nxp;
    load_twiddles(tx, ty, twx, twy, g.Nxp, g.Nyp);
    PairIter it = pair_iter((g.nq + 1) / 2);
    if (it.next >= it.end) return;   // block-uniform
    const float rly = 1.0f / (float)g.Ly, rnyp = 1.0f / (float)g.Nyp;
    const int npx = g.Lx * g.Ly;
    float2 a[kSubUnroll], b[kSubUnroll];
    unsigned cj = 0u;
    sub_bwd_fetch(bt, g, it.next, a, b, cj);
    for (int pair = it.next; pair < it.end; pair += it.step) {
        const int q0 = 2 * pair;
        const bool has1 = q0 + 1 < g.nq;
        const float s0 = qscale ? qscale[q0] : 1.0f;
        const float s1 = (qscale && has1) ? qscale[q0 + 1] : 1.0f;
        lds_barrier();   // the previous pair's grids are consumed (twiddles loaded, first pass)
#pragma unroll
        for (int u = 0; u < kSubUnroll; ++u) {
            const int e = u * kSubThreads + (int)threadIdx.x;
            if (e < npix) {
                const bool c = (cj >> u) & 1u;
                const float2 x = make_float2(a[u].x * s0, (c ? -a[u].y : a[u].y) * s0);
                const float2 y = make_float2(b[u].x * s1, (c ? -b[u].y : b[u].y) * s1);
                const int kx = qdiv(e, g.Nyp, rnyp), ky = e - kx * g.Nyp;
                bufa[kx * P + ky] = make_float2(x.x - y.y, x.y + y.x);   // A + i B
            }
        }
        lds_barrier();
        if (pair + it.step < it.end) sub_bwd_fetch(bt, g, pair + it.step, a, b, cj);
        float2 *r1 = fft_pass_2d(bufa, bufb, dx, tx, 1.0f, g.Nyp, 1, P);                // along x
        float2 *r2 = fft_pass_2d(r1, r1 == bufa ? bufb : bufa, dy, ty, 1.0f, g.Nxp, P, 1);    // along y
        const int c0 = slot_plane[q0 / g.n2], ab0 = q0 % g.n2;
        const int c1 = has1 ? slot_plane[(q0 + 1) / g.n2] : 0, ab1 = has1 ? (q0 + 1) % g.n2 : 0;
        if (has1 && c1 == c0) {   // neighbours in one lenslet record (uniform per pair): one 8-byte store
            for (int e = threadIdx.x; e < npx; e += blockDim.x) {
                const int Qx = qdiv(e, g.Ly, rly), Qy = e - Qx * g.Ly;
                const float2 z = r2[Qx * P + Qy];
                const int64_t cell = ((int64_t)Qx + g.T) * g.YP + (Qy + g.T);
                st_f2_a4(dfft + ((int64_t)c0 * g.XP * g.YP + cell) * g.S + ab0, make_float2(z.x * scale, z.y * scale));
            }
        } else {
            for (int e = threadIdx.x; e < npx; e += blockDim.x) {
                const int Qx = qdiv(e, g.Ly, rly), Qy = e - Qx * g.Ly;
                const float2 z = r2[Qx * P + Qy];
                const int64_t cell = ((int64_t)Qx + g.T) * g.YP + (Qy + g.T);
                dfft[((int64_t)c0 * g.XP * g.YP + cell) * g.S + ab0] = z.x * scale;
                if (has1) dfft[((int64_t)c1 * g.XP * g.YP + cell) * g.S + ab1] = z.y * scale;
            }
        }
    }
}

// Spectra build: kernel grids for a batch of (slot, ab) pairs.
struct SpecJob {
    int plane, ab, q, lo, hi, off;
};

__global__ void k_spec_fill(const float *__restrict__ H, const SpecJob *__restrict__ jobs, int64_t M, int n, int Nx,
                            int Ny, float2 *__restrict__ grids, int nb)
{
    const int64_t per = (int64_t)Nx * Ny;
    const int64_t total = per * nb;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int b = (int)(e / per);
        const int64_t r = e - (int64_t)b * per;
        const int x = (int)(r / Ny), y = (int)(r % Ny);
        const SpecJob jb = jobs[b];
        const int a = jb.ab % n, bb = jb.ab / n;
        int i = (x + jb.off - a) % Nx;
        if (i < 0) i += Nx;
        int jj = (y + jb.off - bb) % Ny;
        if (jj < 0) jj += Ny;
        float v = 0.0f;
        if (i >= jb.lo && i <= jb.hi && jj >= jb.lo && jj <= jb.hi)
            v = H[(int64_t)i + M * ((int64_t)jj + M * ((int64_t)a + n * ((int64_t)bb + (int64_t)n * jb.plane)))];
        grids[e] = make_float2(v, 0.0f);
    }
}

__device__ __forceinline__ void spec_st(float2 *p, float2 v, float) { *p = v; }
__device__ __forceinline__ void spec_st(__half2 *p, float2 v, float inv) { *p = __floats2half2_rn(v.x * inv, v.y * inv); }

template <typename GT>
__global__ void k_spec_scatter(const float2 *__restrict__ grids, const SpecJob *__restrict__ jobs,
                               const float *__restrict__ qscale, int nb, int Nx, int Ny, int Nxp, int Nyp, int Hy,
                               int n, int64_t kstride, GT *__restrict__ G)
{
    const int n2 = n * n;
    const int64_t nk0 = (int64_t)Nxp * Hy;
    const int64_t total = (int64_t)nb * nk0 * n2;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(e % n2);
        const int64_t t = e / n2;
        const int k0 = (int)(t % nk0), b = (int)(t / nk0);
        const int k0x = k0 / Hy, k0y = k0 - k0x * Hy;
        const int kx = k0x + Nxp * (j % n), ky = k0y + Nyp * (j / n);
        const int q = jobs[b].q;
        const float inv = qscale ? 1.0f / qscale[q] : 1.0f;
        spec_st(G + (int64_t)k0 * kstride + (int64_t)q * n2 + j,
                grids[(int64_t)b * Nx * Ny + (int64_t)kx * Ny + ky], inv);
    }
}

// qscale[q] = sum |H| over the kernel box of job q (1 for an all-zero kernel):
// a bound on |G[.][q][.]|, attained at frequency 0 by a non-negative kernel.
__global__ __launch_bounds__(256) void k_spec_l1(const float *__restrict__ H, const SpecJob *__restrict__ jobs,
                                                 int64_t M, int n, float *__restrict__ qscale)
{
    __shared__ double red[256];
    const SpecJob jb = jobs[blockIdx.x];
    const int a = jb.ab % n, bb = jb.ab / n;
    const int side = jb.hi - jb.lo + 1;
    const float *h = H + M * M * ((int64_t)a + n * ((int64_t)bb + (int64_t)n * jb.plane));
    double s = 0.0;
    for (int e = threadIdx.x; e < side * side; e += 256) {
        const int i = jb.lo + e % side, jj = jb.lo + e / side;
        s += fabs((double)h[(int64_t)i + M * jj]);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) qscale[jb.q] = red[0] > 0.0 ? (float)red[0] : 1.0f;
}

static int grid_n(int64_t total) { return (int)std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 65536)); }

void fft_mark(FftState &f, hipStream_t st, int kind, bool begin)
{
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, st);
    f.ev.push_back(e);
    if (begin) f.ev_kind.push_back(kind);
}

// Sum the GEMV times recorded since the last collection (call after a sync).
void fft_timing_collect(sid_psf *p, double *fwd_ms, double *bwd_ms)
{
    FftState &f = p->plan.fft;
    *fwd_ms = *bwd_ms = 0.0;
    for (size_t i = 0; i + 1 < f.ev.size() && i / 2 < f.ev_kind.size(); i += 2) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, f.ev[i], f.ev[i + 1]) == hipSuccess) (f.ev_kind[i / 2] ? *bwd_ms : *fwd_ms) += ms;
    }
    for (auto e : f.ev) (void)hipEventDestroy(e);
    f.ev.clear();
    f.ev_kind.clear();
}

void fft_state_free(FftState &f)
{
    tf_free(f);
    for (auto e : f.ev) (void)hipEventDestroy(e);
    f.ev.clear();
    fft_plan_free(f.fx); fft_plan_free(f.fy); fft_plan_free(f.fX); fft_plan_free(f.fY);
    for (void *q : {(void *)f.G, (void *)f.subs, (void *)f.vt, (void *)f.oh, (void *)f.ohT, (void *)f.grid, (void *)f.dfft,
                    (void *)f.d_slot_plane, (void *)f.d_plane_fft, (void *)f.d_plane_list, (void *)f.qscale})
        if (q) (void)hipFree(q);
    f = FftState();
}

// Planes that can use the FFT path: nonempty with an odd support (the adjoint
// of the FFT forward is then the reference backward, off_b == off_f).
bool fft_eligible(const sid_psf *p, int64_t c)
{
    const Plane &pl = p->planes[c];
    return !pl.empty && pl.off_b == pl.off_f;
}

// Build the FFT geometry and spectra for the planes marked in `use`.
int fft_state_build(sid_psf *p, const std::vector<char> &use, hipStream_t st)
{
    SensorPlan &sp = p->plan;
    FftState &f = sp.fft;
    fft_state_free(f);
    const int n = (int)p->n, n2 = n * n;
    f.slot_of.assign(p->nzH, -1);
    int reach = 0;
    for (int64_t c = 0; c < p->nzH; ++c) {
        if (!use[c]) continue;
        const Plane &pl = p->planes[c];
        f.slot_of[c] = (int)f.planes.size();
        f.planes.push_back((int)c);
        reach = std::max(reach, std::max(pl.hi - pl.off_f, pl.off_f - pl.lo));
    }
    if (f.planes.empty()) return SID_OK;
    if (n2 > 256) {
        set_error("FFT path supports Nnum <= 16");
        fft_state_free(f);
        return SID_ERR_DIMS;
    }
    f.nq = (int)f.planes.size() * n2;
    f.n2 = n2;
    // FP16 spectra in the BF16X2 mode, row q scaled by 1/qscale[q]
    f.half = p->precision == SID_PREC_BF16X2;
    const bool tiled = f.half && std::getenv("SID_FFT_UNTILED") == nullptr;
    if (tiled) {
        // the tiled form groups slots of similar lenslet tap range: ascending width
        std::stable_sort(f.planes.begin(), f.planes.end(), [&](int a, int b) {
            return p->planes[a].tf_hi - p->planes[a].tf_lo < p->planes[b].tf_hi - p->planes[b].tf_lo;
        });
        for (size_t sl = 0; sl < f.planes.size(); ++sl) f.slot_of[f.planes[sl]] = (int)sl;
    }
    std::vector<SpecJob> jobs;
    for (size_t sl = 0; sl < f.planes.size(); ++sl) {
        const Plane &pl = p->planes[f.planes[sl]];
        for (int ab = 0; ab < n2; ++ab)
            jobs.push_back(SpecJob{f.planes[sl], ab, (int)sl * n2 + ab, pl.lo, pl.hi, pl.off_f});
    }
    SpecJob *d_jobs = nullptr;
    SID_TRY_HIP(hipMalloc(&d_jobs, sizeof(SpecJob) * jobs.size()));
    if (hipMemcpy(d_jobs, jobs.data(), sizeof(SpecJob) * jobs.size(), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d_jobs);
        set_error("copy of the spectra jobs failed");
        return SID_ERR_HIP;
    }
    if (f.half) {
        SID_TRY_HIP(hipMalloc(&f.qscale, sizeof(float) * f.nq));
        hipLaunchKernelGGL(k_spec_l1, dim3((unsigned)jobs.size()), dim3(256), 0, st, p->H, d_jobs, p->M, n, f.qscale);
        SID_TRY_HIP(hipGetLastError());
    }
    SID_TRY_HIP(hipMalloc(&f.dfft, sizeof(float) * sp.vol_elems));
    SID_TRY_HIP(hipMemsetAsync(f.dfft, 0, sizeof(float) * sp.vol_elems, st));
    std::vector<int> pf(p->nzH, 0);
    for (int c : f.planes) pf[c] = 1;
    SID_TRY_HIP(hipMalloc(&f.d_slot_plane, sizeof(int) * f.planes.size()));
    SID_TRY_HIP(hipMalloc(&f.d_plane_fft, sizeof(int) * p->nzH));
    SID_TRY_HIP(hipMemcpy(f.d_slot_plane, f.planes.data(), sizeof(int) * f.planes.size(), hipMemcpyHostToDevice));
    SID_TRY_HIP(hipMemcpy(f.d_plane_fft, pf.data(), sizeof(int) * p->nzH, hipMemcpyHostToDevice));
    {
        // the backward reduction of the direct planes runs on the GEMM's stream
        // as soon as the GEMM is done, that of the FFT planes after the FFT path
        std::vector<int> lst;
        for (int64_t c = 0; c < p->nzH; ++c)
            if (!pf[c]) lst.push_back((int)c);
        f.n_list_direct = (int)lst.size();
        for (int64_t c = 0; c < p->nzH; ++c)
            if (pf[c]) lst.push_back((int)c);
        f.n_list_fft = (int)lst.size() - f.n_list_direct;
        f.all_fft = true;
        for (int64_t c = 0; c < p->nzH; ++c)
            if (!pf[c] && !p->planes[c].empty) f.all_fft = false;
        SID_TRY_HIP(hipMalloc(&f.d_plane_list, sizeof(int) * std::max<size_t>(lst.size(), 1)));
        SID_TRY_HIP(hipMemcpy(f.d_plane_list, lst.data(), sizeof(int) * lst.size(), hipMemcpyHostToDevice));
    }
    // BF16X2: the tiled form (SID_FFT_UNTILED=1 keeps the whole-sensor spectra)
    if (tiled) {
        const int rc = tf_build(p, st);
        if (rc == SID_OK) {
            (void)hipFree(d_jobs);
            f.active = true;
            return SID_OK;
        }
        tf_free(f);
        if (rc != kTfNotApplicable) {
            (void)hipFree(d_jobs);
            fft_state_free(f);
            return rc;
        }
    }
    const int Xn = (int)(sp.Lx * n), Yn = (int)(sp.Ly * n);
    f.Nxp = next_smooth235((Xn + reach + n - 1) / n);
    f.Nyp = next_smooth235((Yn + reach + n - 1) / n);
    f.Nx = n * f.Nxp;
    f.Ny = n * f.Nyp;
    if (f.Nx > 2048 || f.Ny > 2048) {
        (void)hipFree(d_jobs);
        set_error("FFT grid %d x %d exceeds the 2048-point transforms of this build", f.Nx, f.Ny);
        fft_state_free(f);
        return SID_ERR_DIMS;
    }
    f.Hy = f.Nyp / 2 + 1;
    f.nk0 = f.Nxp * f.Hy;
    // GEMV row chunks, multiples of 32 rows: about 32K blocks per GEMV (many
    // rounds of the resident blocks, so no half-empty last round).
    // SID_GEMV_SPLIT overrides both counts (measurement).
    {
        auto chunks = [&](int target, int &ns, int &qc) {
            ns = (int)std::max<int64_t>(1, (target + f.nk0 - 1) / f.nk0);
            if (const char *e = std::getenv("SID_GEMV_SPLIT")) ns = std::max(1, std::atoi(e));
            ns = std::min(ns, std::max(1, f.nq / 32));
            qc = ((f.nq + ns - 1) / ns + 31) / 32 * 32;
            ns = (f.nq + qc - 1) / qc;
        };
        chunks(32768, f.nsplit, f.qc);
        chunks(32768, f.nsplit_b, f.qc_b);
    }
    SID_TRY(fft_plan_make(f.fx, f.Nxp));
    SID_TRY(fft_plan_make(f.fy, f.Nyp));
    SID_TRY(fft_plan_make(f.fX, f.Nx));
    SID_TRY(fft_plan_make(f.fY, f.Ny));
    // k0 blocks padded to 16 bytes
    f.kstride = ((int64_t)f.nq * n2 + 3) / 4 * 4;
    f.g_bytes = (int64_t)f.nk0 * f.kstride * (int64_t)(f.half ? sizeof(__half2) : sizeof(float2));
    SID_TRY_HIP(hipMalloc(&f.G, f.g_bytes));
    if ((reinterpret_cast<uintptr_t>(f.G) & 15) != 0) {   // the backward GEMV's 16-byte tile loads rely on it
        set_error("spectra buffer is not 16-byte aligned");
        fft_state_free(f);
        return SID_ERR_HIP;
    }
    // fused sub-image transforms when both Nx' x Ny' grids fit in a block's LDS
    // (SID_FFT_UNFUSED=1 keeps the separate gather / 1-D FFT / pack kernels)
    {
        const size_t lds = sizeof(float2) * (2 * (size_t)f.Nxp * sub_pitch(f.Nyp) + f.Nxp + f.Nyp);
        hipDeviceProp_t prop;
        f.fused = std::getenv("SID_FFT_UNFUSED") == nullptr &&
                  (int64_t)f.Nxp * f.Nyp <= (int64_t)kSubUnroll * kSubThreads &&
                  hipGetDeviceProperties(&prop, p->device) == hipSuccess && lds <= prop.sharedMemPerBlock;
        f.sub_blocks = f.fused ? std::max(1, std::min(prop.multiProcessorCount, (f.nq + 1) / 2)) : 0;
        if (f.fused && lds > 65536) {
            f.fused = hipFuncSetAttribute((const void *)k_fft_sub_fwd, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)lds) == hipSuccess &&
                      hipFuncSetAttribute((const void *)k_fft_sub_bwd, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)lds) == hipSuccess;
            (void)hipGetLastError();
        }
        f.sub_lds = lds;
    }
    if (!f.fused) SID_TRY_HIP(hipMalloc(&f.subs, sizeof(float2) * (int64_t)f.nq * f.Nxp * f.Nyp));
    SID_TRY_HIP(hipMalloc(&f.vt, sizeof(float2) * (int64_t)f.nk0 * f.nq));
    SID_TRY_HIP(hipMalloc(&f.oh, sizeof(float2) * (int64_t)f.nsplit * f.nk0 * n2));
    SID_TRY_HIP(hipMalloc(&f.ohT, sizeof(float2) * (int64_t)f.nk0 * n2));
    SID_TRY_HIP(hipMalloc(&f.grid, sizeof(float2) * (int64_t)f.Nx * f.Ny));

    // spectra, in batches of kernels
    const int64_t per = (int64_t)f.Nx * f.Ny;
    const int bmax = (int)std::max<int64_t>(1, std::min<int64_t>(64, (int64_t)(512ll << 20) / (per * 8)));
    float2 *grids = nullptr;
    SID_TRY_HIP(hipMalloc(&grids, sizeof(float2) * per * bmax));
    int rc = SID_OK;
    for (size_t b0 = 0; b0 < jobs.size() && rc == SID_OK; b0 += bmax) {
        const int nb = (int)std::min<size_t>(bmax, jobs.size() - b0);
        hipLaunchKernelGGL(k_spec_fill, dim3(grid_n(per * nb)), dim3(256), 0, st, p->H, d_jobs + b0, p->M, n, f.Nx,
                           f.Ny, grids, nb);
        // 2-D forward DFT of each grid: along y (contiguous), then along x
        rc = fft_run(f.fY, grids, (int64_t)nb * f.Nx, (int64_t)nb * f.Nx, 0, f.Ny, 1, -1, 1.0f, st);
        if (rc == SID_OK) rc = fft_run(f.fX, grids, (int64_t)nb * f.Ny, f.Ny, per, 1, f.Ny, -1, 1.0f, st);
        if (rc == SID_OK) {
            const dim3 gs(grid_n((int64_t)nb * f.nk0 * n2));
            if (f.half)
                hipLaunchKernelGGL(k_spec_scatter<__half2>, gs, dim3(256), 0, st, grids, d_jobs + b0, f.qscale, nb,
                                   f.Nx, f.Ny, f.Nxp, f.Nyp, f.Hy, n, f.kstride, (__half2 *)f.G);
            else
                hipLaunchKernelGGL(k_spec_scatter<float2>, gs, dim3(256), 0, st, grids, d_jobs + b0,
                                   (const float *)nullptr, nb, f.Nx, f.Ny, f.Nxp, f.Nyp, f.Hy, n, f.kstride,
                                   (float2 *)f.G);
        }
    }
    const hipError_t e1 = hipStreamSynchronize(st);
    (void)hipFree(grids);
    (void)hipFree(d_jobs);
    if (rc != SID_OK) return rc;
    if (e1 != hipSuccess) {
        set_error("FFT spectra build failed: %s", hipGetErrorString(e1));
        return SID_ERR_HIP;
    }
    f.active = true;
    return SID_OK;
}

static SubGeom sub_geom(const sid_psf *p, const FftState &f)
{
    const SensorPlan &sp = p->plan;
    SubGeom g;
    g.nq = f.nq;
    g.n2 = f.n2;
    g.Nxp = f.Nxp;
    g.Nyp = f.Nyp;
    g.Hy = f.Hy;
    g.Lx = (int)sp.Lx;
    g.Ly = (int)sp.Ly;
    g.T = (int)p->T;
    g.S = (int)p->S;
    g.nz_in = (int)sp.nz;
    g.XP = sp.XP;
    g.YP = sp.YP;
    return g;
}

// Forward FFT planes: leaves the real N x N image in f.grid (.x).
int fft_forward(sid_psf *p, hipStream_t st)
{
    SensorPlan &sp = p->plan;
    FftState &f = sp.fft;
    if (!f.active) return SID_OK;
    if (f.tiled) return tf_forward(p, st);
    const int n2 = f.n2;
    if (f.fused) {
        hipLaunchKernelGGL(k_fft_sub_fwd, dim3(f.sub_blocks), dim3(kSubThreads), f.sub_lds, st, sp.vol, f.d_slot_plane,
                           f.qscale, sub_geom(p, f), fft_desc(f.fx), fft_desc(f.fy), f.fx.tw_fwd, f.fy.tw_fwd, f.vt);
        SID_TRY_HIP(hipGetLastError());
    } else {
        hipLaunchKernelGGL(k_fft_sub_gather, dim3(grid_n((int64_t)f.nq * f.Nxp * f.Nyp)), dim3(256), 0, st, sp.vol,
                           f.d_slot_plane, f.nq, n2, f.Nxp, f.Nyp, (int)sp.Lx, (int)sp.Ly, sp.XP, sp.YP, (int)p->T,
                           (int)p->S, (int)sp.nz, f.subs);
        SID_TRY_HIP(hipGetLastError());
        const int64_t sz = (int64_t)f.Nxp * f.Nyp;
        SID_TRY(fft_run(f.fy, f.subs, (int64_t)f.nq * f.Nxp, (int64_t)f.nq * f.Nxp, 0, f.Nyp, 1, -1, 1.0f, st));
        SID_TRY(fft_run(f.fx, f.subs, (int64_t)f.nq * f.Hy, f.Hy, sz, 1, f.Nyp, -1, 1.0f, st));
        hipLaunchKernelGGL(k_fft_pack_vt, dim3(grid_n((int64_t)f.nk0 * f.nq)), dim3(256), 0, st, f.subs, f.qscale, f.nq,
                           f.Nxp, f.Nyp, f.Hy, f.vt);
    }
    fft_mark(f, st, 0, true);
    if (f.half)
        hipLaunchKernelGGL(k_fft_gemv_fwd<__half2>, dim3(f.nk0, f.nsplit), dim3(256), 0, st, (const __half2 *)f.G,
                           f.vt, f.nq, n2, f.qc, f.nk0, f.kstride, f.oh);
    else
        hipLaunchKernelGGL(k_fft_gemv_fwd<float2>, dim3(f.nk0, f.nsplit), dim3(256), 0, st, (const float2 *)f.G,
                           f.vt, f.nq, n2, f.qc, f.nk0, f.kstride, f.oh);
    fft_mark(f, st, 0, false);
    hipLaunchKernelGGL(k_fft_oh_sum, dim3((f.nk0 + 31) / 32, (n2 + 31) / 32), dim3(256), 0, st, f.oh, f.nsplit, f.nk0,
                       n2, f.ohT);
    hipLaunchKernelGGL(k_fft_assemble, dim3(grid_n((int64_t)f.Nx * f.Ny)), dim3(256), 0, st, f.ohT, f.nk0, f.Nx, f.Ny,
                       f.Nxp, f.Nyp, f.Hy, (int)p->n, f.grid);
    SID_TRY_HIP(hipGetLastError());
    SID_TRY(fft_run(f.fY, f.grid, f.Nx, f.Nx, 0, f.Ny, 1, +1, 1.0f, st));
    SID_TRY(fft_run(f.fX, f.grid, f.Ny, f.Ny, 0, 1, f.Ny, +1, 1.0f / ((float)f.Nx * (float)f.Ny), st));
    return SID_OK;
}

// Backward FFT planes from the image in sp.img: writes f.dfft (padded volume layout).
int fft_backward(sid_psf *p, hipStream_t st, const BwdTail *tail, bool *tail_done)
{
    SensorPlan &sp = p->plan;
    FftState &f = sp.fft;
    if (tail_done) *tail_done = false;
    if (!f.active) return SID_OK;
    if (f.tiled) return tf_backward(p, st, tail, tail_done);
    const int n = (int)p->n;
    hipLaunchKernelGGL(k_fft_img_grid, dim3(grid_n((int64_t)f.Nx * f.Ny)), dim3(256), 0, st, sp.img, f.Nx, f.Ny,
                       (int)(sp.Lx * n), (int)(sp.Ly * n), n, sp.YP, (int)p->T, (int)p->S, f.grid);
    SID_TRY_HIP(hipGetLastError());
    SID_TRY(fft_run(f.fY, f.grid, f.Nx, f.Nx, 0, f.Ny, 1, -1, 1.0f, st));
    SID_TRY(fft_run(f.fX, f.grid, f.Ny, f.Ny, 0, 1, f.Ny, -1, 1.0f, st));
    fft_mark(f, st, 1, true);
    if (f.half)
        hipLaunchKernelGGL(k_fft_gemv_bwd<__half2>, dim3(f.nk0, f.nsplit_b), dim3(256), 0, st, (const __half2 *)f.G,
                           f.grid, f.nq, f.qc_b, f.kstride, n, f.Nx, f.Ny, f.Nxp, f.Nyp, f.Hy, f.vt);
    else
        hipLaunchKernelGGL(k_fft_gemv_bwd<float2>, dim3(f.nk0, f.nsplit_b), dim3(256), 0, st, (const float2 *)f.G,
                           f.grid, f.nq, f.qc_b, f.kstride, n, f.Nx, f.Ny, f.Nxp, f.Nyp, f.Hy, f.vt);
    fft_mark(f, st, 1, false);
    if (f.fused) {
        hipLaunchKernelGGL(k_fft_sub_bwd, dim3(f.sub_blocks), dim3(kSubThreads), f.sub_lds, st, f.vt, f.d_slot_plane,
                           f.qscale, sub_geom(p, f), fft_desc(f.fx), fft_desc(f.fy), f.fx.tw_inv, f.fy.tw_inv,
                           1.0f / ((float)f.n2 * (float)f.Nxp * (float)f.Nyp), f.dfft);
        SID_TRY_HIP(hipGetLastError());
        return SID_OK;
    }
    hipLaunchKernelGGL(k_fft_unpack_bt, dim3(grid_n((int64_t)f.nk0 * f.nq)), dim3(256), 0, st, f.vt, f.qscale, f.nq,
                       f.Nxp, f.Nyp, f.Hy, f.subs);
    const int nhi = f.Nyp - f.Hy;
    if (nhi > 0)
        hipLaunchKernelGGL(k_fft_herm_expand, dim3(grid_n((int64_t)f.nq * f.Nxp * nhi)), dim3(256), 0, st, f.subs,
                           f.nq, f.Nxp, f.Nyp, f.Hy);
    SID_TRY_HIP(hipGetLastError());
    const int64_t sz = (int64_t)f.Nxp * f.Nyp;
    // inverse 2-D sub transforms: along x (stride Nyp), then along y; scale 1 / (n^2 Nx' Ny')
    SID_TRY(fft_run(f.fx, f.subs, (int64_t)f.nq * f.Nyp, f.Nyp, sz, 1, f.Nyp, +1, 1.0f, st));
    SID_TRY(fft_run(f.fy, f.subs, (int64_t)f.nq * f.Nxp, (int64_t)f.nq * f.Nxp, 0, f.Nyp, 1, +1,
                    1.0f / ((float)f.n2 * (float)f.Nxp * (float)f.Nyp), st));
    hipLaunchKernelGGL(k_fft_sub_scatter, dim3(grid_n((int64_t)f.nq * sp.Lx * sp.Ly)), dim3(256), 0, st, f.subs,
                       f.d_slot_plane, f.nq, f.n2, f.Nxp, f.Nyp, (int)sp.Lx, (int)sp.Ly, sp.XP, sp.YP, (int)p->T,
                       (int)p->S, f.dfft);
    SID_TRY_HIP(hipGetLastError());
    return SID_OK;
}

// Per-plane direct-vs-FFT choice from measurements on this device: time the
// forward with every plane direct and with every eligible plane on the FFT
// path, model a plane's direct cost as its share of the MFMA work and its FFT
// cost as its share of the FFT-path time (the spectra stream dominates and
// grows with the number of FFT planes), and put a plane on the FFT path when
// that is cheaper.  Cached per (X, Y, nz, precision).
static double time_ms(hipStream_t st, int reps, const std::function<int()> &fn, int *rc)
{
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) { *rc = SID_ERR_HIP; return 0; }
    *rc = fn();  // warm-up
    if (*rc == SID_OK) (void)hipEventRecord(a, st);
    for (int r = 0; r < reps && *rc == SID_OK; ++r) *rc = fn();
    float ms = 0.f;
    if (*rc == SID_OK) {
        (void)hipEventRecord(b, st);
        (void)hipEventSynchronize(b);
        (void)hipEventElapsedTime(&ms, a, b);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms / reps;
}

int autotune_methods(sid_psf *p, int64_t X, int64_t Y, int64_t nz, hipStream_t st)
{
    if (p->method != SID_METHOD_AUTO || p->precision == SID_PREC_FP64) return SID_OK;
    if (p->pin.size() == (size_t)p->nzH && std::find(p->pin.begin(), p->pin.end(), 0) == p->pin.end())
        return SID_OK;  // every plane pinned: nothing to choose
    const int64_t key[4] = {X, Y, nz, p->precision};
    if (std::equal(key, key + 4, p->auto_key) && p->auto_fft.size() == (size_t)p->nzH) return SID_OK;
    std::vector<char> elig(p->nzH, 0);
    int ne = 0;
    for (int64_t c = 0; c < p->nzH; ++c) {
        elig[c] = fft_eligible(p, c) ? 1 : 0;
        ne += elig[c];
    }
    p->auto_fft.assign(p->nzH, 0);
    p->auto_refined = false;
    std::copy(key, key + 4, p->auto_key);
    if (ne == 0) return SID_OK;
    if (p->precision == SID_PREC_BF16X2 && std::getenv("SID_FFT_UNTILED") == nullptr && p->n * p->n <= 256) {
        // BF16X2 with the tiled FFT path: every eligible plane goes there, untimed.
        // Its FP16 spectra (2^-11) perturb the operator far less than the BF16 PSF
        // (2^-9) of direct planes, and the ISRA loop accumulates that perturbation:
        // at 1005^2 x 51, 8 iterations, rel-L2 vs FP64 1.1e-4 with every plane on the
        // FFT path against 7.6e-4 with the 6 narrowest planes direct and 8.0e-4 with
        // 11, for 2-3 % more speed.  Pinned planes keep their pins.
        for (int64_t c = 0; c < p->nzH; ++c) p->auto_fft[c] = elig[c];
        p->auto_refined = true;
        p->auto_ms[0] = p->auto_ms[1] = 0.0;
        p->auto_ms[2] = p->auto_ms[3] = ne;
        return SID_OK;
    }
    int rc = SID_OK;
    // 1. every plane direct
    p->method = SID_METHOD_DIRECT;
    rc = ensure_plan(p, X, Y, nz);
    double t_dir = 0;
    std::vector<double> macs(p->nzH, 0.0);
    double macs_all = 0;
    if (rc == SID_OK) {
        t_dir = time_ms(st, 2, [&]() { return launch_gemm(p, 0, st); }, &rc);
        const GemmPlan &F = p->plan.fwd;
        for (size_t s = 0; s < F.segs.size(); ++s) {
            const double m = (double)F.cost[s] * p->NP * p->plan.mpad;
            macs[F.segs[s].plane] += m;
            macs_all += m;
        }
    }
    // 2. every eligible plane on the FFT path
    double t_fft = 0;
    if (rc == SID_OK) {
        p->method = SID_METHOD_FFT;
        rc = ensure_plan(p, X, Y, nz);
        if (rc == SID_OK) t_fft = time_ms(st, 2, [&]() { return fft_forward(p, st); }, &rc);
    }
    p->method = SID_METHOD_AUTO;
    if (rc != SID_OK) {
        p->auto_key[0] = -1;
        return rc;
    }
    const double fft_per_plane = t_fft / ne;
    int chosen = 0;
    for (int64_t c = 0; c < p->nzH; ++c) {
        const double direct_ms = macs_all > 0 ? t_dir * macs[c] / macs_all : 0.0;
        p->auto_fft[c] = (elig[c] && fft_per_plane < direct_ms) ? 1 : 0;
        chosen += p->auto_fft[c];
    }
    p->auto_ms[0] = t_dir;
    p->auto_ms[1] = t_fft;
    p->auto_ms[2] = ne;
    p->auto_ms[3] = chosen;
    return SID_OK;
}

int fft_test(int L, int64_t nseq, float2 *data, int dir, hipStream_t st)
{
    FftPlan1D pl;
    SID_TRY(fft_plan_make(pl, L));
    const int rc = fft_run(pl, data, nseq, nseq, 0, L, 1, dir, 1.0f, st);
    const hipError_t e = hipStreamSynchronize(st);
    fft_plan_free(pl);
    if (rc != SID_OK) return rc;
    if (e != hipSuccess) {
        set_error("FFT test failed: %s", hipGetErrorString(e));
        return SID_ERR_HIP;
    }
    return SID_OK;
}

}  // namespace sid
