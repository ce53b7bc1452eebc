// convn(V, k, 'same') on column-major device volumes: the neuron-kernel
// wrapping of the projectors in reconstruct_S.m:122-128,184-190 and
// generate_LFM_library_GPU.m:95-101 (forwardFUN = F(convn(X, k, 'same')),
// backwardFUN = convn(B(y), k, 'same')).
//
// MATLAB's 'same' window of the full convolution starts at floor(size(k)/2):
//   out[x, y, z] = sum_{i,j,l} V[x + ox - i, y + oy - j, z + oz - l] k[i, j, l],
//   (ox, oy, oz) = floor((kx, ky, kz) / 2), V = 0 outside the volume.
//
// The direct forms here run where the transform form (sid_convn_fft.hip) does
// not apply or SID_CONVN_FFT=0.  FP32 / FP64: k_convn_same, the FP32 direct
// sum.  One block computes a 64 x 16 (x, y) output tile of one plane.  For
// every kernel slice l the (64 + kx - 1) x (16 + ky - 1) input tile of plane
// z + oz - l is staged in LDS; a thread owns 4 consecutive x outputs and walks
// a 4-wide register window along the kernel row, so each LDS read feeds four
// FMAs.  The slice's taps sit in LDS beside the tile (broadcast reads).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "sid_internal.h"

namespace sid {

constexpr int kCvTx = 64, kCvTy = 16;

__global__ __launch_bounds__(256) void k_convn_same(const float *__restrict__ V, int X, int Y, int Z,
                                                    const float *__restrict__ k, int kx, int ky, int kz,
                                                    float *__restrict__ out)
{
    extern __shared__ float tile[];
    const int WX = kCvTx + kx - 1, WY = kCvTy + ky - 1;
    float *ks = tile + WX * WY;   // taps of the current slice, [j][i]
    const int x0 = blockIdx.x * kCvTx, y0 = blockIdx.y * kCvTy, z = blockIdx.z;
    const int ox = kx / 2, oy = ky / 2, oz = kz / 2;
    const int tx = (threadIdx.x & 15) * 4, ty = threadIdx.x >> 4;   // 16 x 16 threads, 4 x outputs each
    // tile element (u, w) holds V[x0 + ox - (kx - 1) + u, y0 + oy - (ky - 1) + w]; output (tx + m, ty)
    // with tap (i, j) reads tile (tx + m + kx - 1 - i, ty + ky - 1 - j)
    const int bx = x0 + ox - (kx - 1), by = y0 + oy - (ky - 1);
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int l = 0; l < kz; ++l) {
        const int zi = z + oz - l;
        if (zi < 0 || zi >= Z) continue;   // block-uniform
        const float *src = V + (int64_t)zi * X * Y;
        __syncthreads();
        for (int base = 0; base < WX * WY; base += 8 * 256) {
            float v[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int e = base + q * 256 + (int)threadIdx.x;
                const int w = e / WX, u = e - w * WX;
                const int xs = bx + u, ys = by + w;
                v[q] = (e < WX * WY && xs >= 0 && xs < X && ys >= 0 && ys < Y) ? src[(int64_t)ys * X + xs] : 0.0f;
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int e = base + q * 256 + (int)threadIdx.x;
                if (e < WX * WY) tile[e] = v[q];
            }
        }
        for (int e = threadIdx.x; e < kx * ky; e += 256) ks[e] = k[(int64_t)l * kx * ky + e];
        __syncthreads();
        for (int j = 0; j < ky; ++j) {
            const float *row = tile + (ty + ky - 1 - j) * WX + tx + kx - 1;   // row[m - i] for output tx + m
            const float *kr = ks + j * kx;
            float w0 = row[0], w1 = row[1], w2 = row[2], w3 = row[3];
            for (int i = 0; i < kx; ++i) {
                const float kv = kr[i];
                acc[0] += w0 * kv;
                acc[1] += w1 * kv;
                acc[2] += w2 * kv;
                acc[3] += w3 * kv;
                // slide: next i reads row[m - i - 1]
                w3 = w2;
                w2 = w1;
                w1 = w0;
                w0 = i + 1 < kx ? row[-(i + 1)] : 0.0f;
            }
        }
    }
    const int y = y0 + ty;
    if (y >= Y) return;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int x = x0 + tx + m;
        if (x < X) out[((int64_t)z * Y + y) * X + x] = acc[m];
    }
}

// ---------------------------------------------------------------------------
// BF16X2 / BF16X3 beyond the transform form's lengths (sid_convn_fft.hip) or
// with SID_CONVN_FFT=0: the same sum on MFMA (v_mfma_f32_16x16x32_bf16).
// For one kernel row (j, l) the convolution along x of 16 output columns is a
// GEMM with a 32 x 16 Toeplitz matrix of that row:
//   out[x0 + m, y, z] += sum_t In[x0 + ox - (kx - 1) + t, y + oy - j, z + oz - l] * B[t][m],
//   B[t][m] = k[m + kx - 1 - t, j, l] (0 outside the kernel),  t in [0, 32 KS), KS = ceil((15 + kx) / 32).
// MFMA rows are 16 output rows y, columns 16 output x, K the input x window.
// A block computes a 64 x 64 output tile of one plane (4 waves x 16 rows, 4
// x-blocks of 16 per wave).  Per kernel slice l the input tile (with halo) is
// staged in LDS as BF16 hi and lo planes (zero outside the volume) and the
// precomputed B fragments of the slice's rows after it.  Like the projector's
// precision modes: BF16X2 = A_hi B_hi + A_lo B_hi (kernel quantised to BF16),
// BF16X3 adds A_hi B_lo.
// ---------------------------------------------------------------------------

// B fragments: frag[(((l * ky + j) * KS + ks) * 2 + part) * 64 + lane][s], lane = m + 16 g holds
// B[32 ks + 8 g + s][m] as BF16 hi (part 0) and lo (part 1).
__global__ void k_convn_bfrag(const float *__restrict__ k, int kx, int ky, int kz, int KS, u16 *__restrict__ frag)
{
    const int64_t total = (int64_t)kz * ky * KS * 64;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int lane = (int)(e & 63);
        int64_t rest = e >> 6;
        const int ks = (int)(rest % KS);
        rest /= KS;
        const int j = (int)(rest % ky), l = (int)(rest / ky);
        const int m = lane & 15, g = lane >> 4;
        u16 *hi = frag + ((((int64_t)l * ky + j) * KS + ks) * 2 * 64 + lane) * 8;
        u16 *lo = hi + 64 * 8;
        for (int s = 0; s < 8; ++s) {
            const int t = 32 * ks + 8 * g + s, i = m + kx - 1 - t;
            const float v = (i >= 0 && i < kx) ? k[i + (int64_t)kx * (j + (int64_t)ky * l)] : 0.0f;
            const u16 h = f2bf(v);
            hi[s] = h;
            lo[s] = f2bf(v - bf2f(h));
        }
    }
}

constexpr int kCmTile = 64;   // output tile side (x and y) of k_convn_mfma
constexpr int kStageU = 8;    // staging loads in flight per thread

typedef int i32x4 __attribute__((ext_vector_type(4)));

// Lane r of each 16-lane row takes lane r - 1's value (DPP row_shr:1); lane 0 keeps its own.
__device__ __forceinline__ bf8 shift_rows(bf8 v)
{
    i32x4 x = __builtin_bit_cast(i32x4, v);
    x.x = __builtin_amdgcn_update_dpp(x.x, x.x, 0x111, 0xf, 0xf, false);
    x.y = __builtin_amdgcn_update_dpp(x.y, x.y, 0x111, 0xf, 0xf, false);
    x.z = __builtin_amdgcn_update_dpp(x.z, x.z, 0x111, 0xf, 0xf, false);
    x.w = __builtin_amdgcn_update_dpp(x.w, x.w, 0x111, 0xf, 0xf, false);
    return __builtin_bit_cast(bf8, x);
}

// The A operand slides with the kernel row: lane r (output row r of the wave)
// reads input tile row w = 16 wave + r + ky - 1 - j, which at step j + 1 is the
// row lane r - 1 held at step j.  So after the first row of a slice only lane 0
// of each 16-lane row reads LDS; the others take their neighbour's registers.
constexpr int kCmPre = 32;    // input values per thread a prefetched slice may hold (tile <= 8192)

// Global loads of one staged input tile (Hh x W values at (X0, Y0) of the slice
// src, zero outside the volume) into registers, kCmPre per thread.
__device__ __forceinline__ void convn_tile_load(float (&v)[kCmPre], const float *__restrict__ src, int n, int W,
                                                int X0, int Y0, int X, int Y)
{
    // an opaque copy of the thread index keeps the per-value index math here
    // instead of hoisted out of the slice loop (where it would hold ~100 registers)
    int tid = (int)threadIdx.x;
    asm volatile("" : "+v"(tid));
#pragma unroll
    for (int q = 0; q < kCmPre; ++q) {
        const int e = q * 256 + tid;
        const int w = e / W, u = e - w * W;
        const int xs = X0 + u, ys = Y0 + w;
        v[q] = (e < n && xs >= 0 && xs < X && ys >= 0 && ys < Y) ? src[(int64_t)ys * X + xs] : 0.0f;
    }
}

// ... and from registers into the BF16 hi / lo tiles (pitch PW).
__device__ __forceinline__ void convn_tile_store(const float (&v)[kCmPre], u16 *thi, u16 *tlo, int n, int W, int PW)
{
    int tid = (int)threadIdx.x;
    asm volatile("" : "+v"(tid));
#pragma unroll
    for (int q = 0; q < kCmPre; ++q) {
        const int e = q * 256 + tid;
        if (e < n) {
            const int w = e / W, u = e - w * W;
            const u16 h = f2bf(v[q]);
            thi[w * PW + u] = h;
            tlo[w * PW + u] = f2bf(v[q] - bf2f(h));
        }
    }
}

template <int MODE, int KS>
__global__ __launch_bounds__(256) void k_convn_mfma(const float *__restrict__ In, int X, int Y, int Z,
                                                    const u16 *__restrict__ frag, int kx, int ky, int kz, int JC,
                                                    float *__restrict__ out)
{
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int W = 48 + 32 * KS;   // input tile width: 3 x-block offsets + the K window
    constexpr int PW = W + 8;         // pitch in BF16 elements (multiple of 8: 16-byte rows)
    constexpr int NBP = MODE == 2 ? 2 : 1;   // B parts staged: hi (and lo for BF16X3)
    const int Hh = kCmTile + ky - 1;  // input tile height
    u16 *thi = reinterpret_cast<u16 *>(smem);
    u16 *tlo = thi + Hh * PW;
    bf8 *bs = reinterpret_cast<bf8 *>(tlo + Hh * PW);   // [jj][ks][part < NBP][64 lanes]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, l16 = lane & 15, g = lane >> 4;
    const int xb0 = blockIdx.x * kCmTile, yb0 = blockIdx.y * kCmTile, z = blockIdx.z;
    const int ox = kx / 2, oy = ky / 2, oz = kz / 2;
    const int X0 = xb0 + ox - (kx - 1), Y0 = yb0 + oy - (ky - 1);
    f4 acc[4];
#pragma unroll
    for (int xb = 0; xb < 4; ++xb) acc[xb] = f4{0.f, 0.f, 0.f, 0.f};
    bf8 ah[4][KS], al[4][KS];
    // slices l with z + oz - l inside the volume: one contiguous range
    const int lb = max(0, z + oz - (Z - 1)), le = min(kz - 1, z + oz);
    // tiles of at most kCmPre values per thread are prefetched into registers one
    // slice ahead (their loads land under the previous slice's MFMAs)
    const bool prefetch = Hh * W <= 256 * kCmPre;
    float nxt[kCmPre];
    if (prefetch && lb <= le) convn_tile_load(nxt, In + (int64_t)(z + oz - lb) * X * Y, Hh * W, W, X0, Y0, X, Y);
    for (int l = lb; l <= le; ++l) {
        const int zi = z + oz - l;
        __syncthreads();                    // tiles of the previous slice consumed
        const float *src = In + (int64_t)zi * X * Y;
        if (prefetch) convn_tile_store(nxt, thi, tlo, Hh * W, W, PW);
        // staging: kStageU independent loads in flight per thread, then convert and store
        for (int base = 0; !prefetch && base < Hh * W; base += kStageU * 256) {
            float v[kStageU];
#pragma unroll
            for (int q = 0; q < kStageU; ++q) {
                const int e = base + q * 256 + (int)threadIdx.x;
                const int w = e / W, u = e - w * W;
                const int xs = X0 + u, ys = Y0 + w;
                v[q] = (e < Hh * W && xs >= 0 && xs < X && ys >= 0 && ys < Y) ? src[(int64_t)ys * X + xs] : 0.0f;
            }
#pragma unroll
            for (int q = 0; q < kStageU; ++q) {
                const int e = base + q * 256 + (int)threadIdx.x;
                if (e < Hh * W) {
                    const int w = e / W, u = e - w * W;
                    const u16 h = f2bf(v[q]);
                    thi[w * PW + u] = h;
                    tlo[w * PW + u] = f2bf(v[q] - bf2f(h));
                }
            }
        }
        for (int jc0 = 0; jc0 < ky; jc0 += JC) {
            const int jn = min(JC, ky - jc0);
            __syncthreads();   // tile staged (first chunk) / previous B chunk consumed
            // fragments are stored [l][j][ks][hi, lo][64 lanes]; BF16X2 stages the hi parts only
            const bf8 *fsrc = reinterpret_cast<const bf8 *>(frag) + ((int64_t)l * ky + jc0) * KS * 2 * 64;
            const int nb = jn * KS * NBP * 64;
            for (int base = 0; base < nb; base += kStageU * 256) {
                bf8 t[kStageU];
#pragma unroll
                for (int q = 0; q < kStageU; ++q) {
                    const int e = base + q * 256 + (int)threadIdx.x;
                    if (e < nb) t[q] = NBP == 2 ? fsrc[e] : fsrc[(e >> 6) * 128 + (e & 63)];
                }
#pragma unroll
                for (int q = 0; q < kStageU; ++q) {
                    const int e = base + q * 256 + (int)threadIdx.x;
                    if (e < nb) bs[e] = t[q];
                }
            }
            // the next slice's input, behind this slice's fragment loads (waits are in issue order)
            if (prefetch && jc0 == 0 && l < le)
                convn_tile_load(nxt, In + (int64_t)(zi - 1) * X * Y, Hh * W, W, X0, Y0, X, Y);
            __syncthreads();
            if (jc0 == 0) {   // all lanes read their row for j = 0
                const int w = wave * 16 + l16 + ky - 1;
#pragma unroll
                for (int xb = 0; xb < 4; ++xb)
#pragma unroll
                    for (int ks = 0; ks < KS; ++ks) {
                        const int u = 16 * xb + 32 * ks + 8 * g;
                        ah[xb][ks] = *reinterpret_cast<const bf8 *>(thi + w * PW + u);
                        al[xb][ks] = *reinterpret_cast<const bf8 *>(tlo + w * PW + u);
                    }
            }
            for (int jj = 0; jj < jn; ++jj) {
                const int j = jc0 + jj;
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) {
                    const bf8 bh = bs[((jj * KS + ks) * NBP + 0) * 64 + lane];
                    bf8 bl;
                    if constexpr (MODE == 2) bl = bs[((jj * KS + ks) * NBP + 1) * 64 + lane];
#pragma unroll
                    for (int xb = 0; xb < 4; ++xb) {
                        f4 c = acc[xb];
                        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[xb][ks], bh, c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[xb][ks], bh, c, 0, 0, 0);
                        if constexpr (MODE == 2)
                            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[xb][ks], bl, c, 0, 0, 0);
                        acc[xb] = c;
                    }
                }
                // slide to row j + 1 (preparing it before row j's MFMAs measured 25 % slower)
                if (j + 1 < ky) {
                    const int w0 = wave * 16 + ky - 2 - j;   // lane 0's new row
#pragma unroll
                    for (int xb = 0; xb < 4; ++xb)
#pragma unroll
                        for (int ks = 0; ks < KS; ++ks) {
                            ah[xb][ks] = shift_rows(ah[xb][ks]);
                            al[xb][ks] = shift_rows(al[xb][ks]);
                        }
                    if (l16 == 0) {
#pragma unroll
                        for (int xb = 0; xb < 4; ++xb)
#pragma unroll
                            for (int ks = 0; ks < KS; ++ks) {
                                const int u = 16 * xb + 32 * ks + 8 * g;
                                ah[xb][ks] = *reinterpret_cast<const bf8 *>(thi + w0 * PW + u);
                                al[xb][ks] = *reinterpret_cast<const bf8 *>(tlo + w0 * PW + u);
                            }
                    }
                }
            }
        }
    }
    // C layout: lane (l16, g) holds C[4 g + i][l16] = out[xb0 + 16 xb + l16, yb0 + 16 wave + 4 g + i]
#pragma unroll
    for (int xb = 0; xb < 4; ++xb) {
        const int x = xb0 + 16 * xb + l16;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int y = yb0 + 16 * wave + 4 * g + i;
            if (x < X && y < Y) out[((int64_t)z * Y + y) * X + x] = acc[xb][i];
        }
    }
}

int convn_ks(int kx) { return (15 + kx + 31) / 32; }

template <int MODE, int KS>
static int launch_convn_mfma(dim3 grid, size_t lds, hipStream_t st, const float *in, const int64_t vdims[3],
                             const u16 *frag, const int kd[3], int JC, float *out)
{
    SID_TRY_HIP(hipFuncSetAttribute((const void *)k_convn_mfma<MODE, KS>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
    hipLaunchKernelGGL((k_convn_mfma<MODE, KS>), grid, dim3(256), lds, st, in, (int)vdims[0], (int)vdims[1],
                       (int)vdims[2], frag, kd[0], kd[1], kd[2], JC, out);
    return SID_OK;
}

int convn_build_frag(const float *k_dev, const int kd[3], u16 **frag, hipStream_t st)
{
    const int KS = convn_ks(kd[0]);
    const int64_t n = (int64_t)kd[2] * kd[1] * KS * 2 * 64 * 8;
    SID_TRY_HIP(hipMalloc(frag, sizeof(u16) * n));
    const int64_t total = (int64_t)kd[2] * kd[1] * KS * 64;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 4096));
    hipLaunchKernelGGL(k_convn_bfrag, dim3(grid), dim3(256), 0, st, k_dev, kd[0], kd[1], kd[2], KS, *frag);
    SID_TRY_HIP(hipGetLastError());
    return SID_OK;
}

int convn_mfma(const float *in, const int64_t vdims[3], const u16 *frag, const int kd[3], int precision, float *out,
               hipStream_t st)
{
    if (vdims[0] < 1 || vdims[1] < 1 || vdims[2] < 1 || vdims[0] > (1 << 24) || vdims[1] > (1 << 24) ||
        vdims[2] > 65535) {
        set_error("convn: volume dimensions out of range");
        return SID_ERR_DIMS;
    }
    const int KS = convn_ks(kd[0]);
    const size_t tile = 4 * (size_t)(kCmTile + kd[1] - 1) * (48 + 32 * KS + 8);
    const size_t per_row = (size_t)KS * (precision == SID_PREC_BF16X3 ? 2 : 1) * 64 * 16;   // staged B parts
    const size_t budget = 150 * 1024;
    const int JC = (int)std::max<size_t>(1, std::min<size_t>(kd[1], (budget - tile) / per_row));
    const size_t lds = tile + JC * per_row;
    const dim3 grid((unsigned)((vdims[0] + kCmTile - 1) / kCmTile), (unsigned)((vdims[1] + kCmTile - 1) / kCmTile),
                    (unsigned)vdims[2]);
    const int mode = precision == SID_PREC_BF16X3 ? 2 : 1;
    if (mode == 2 && KS == 1) SID_TRY((launch_convn_mfma<2, 1>(grid, lds, st, in, vdims, frag, kd, JC, out)));
    else if (mode == 2 && KS == 2) SID_TRY((launch_convn_mfma<2, 2>(grid, lds, st, in, vdims, frag, kd, JC, out)));
    else if (mode == 2) SID_TRY((launch_convn_mfma<2, 3>(grid, lds, st, in, vdims, frag, kd, JC, out)));
    else if (KS == 1) SID_TRY((launch_convn_mfma<1, 1>(grid, lds, st, in, vdims, frag, kd, JC, out)));
    else if (KS == 2) SID_TRY((launch_convn_mfma<1, 2>(grid, lds, st, in, vdims, frag, kd, JC, out)));
    else SID_TRY((launch_convn_mfma<1, 3>(grid, lds, st, in, vdims, frag, kd, JC, out)));
    SID_TRY_HIP(hipGetLastError());
    return SID_OK;
}

int convn_same(const float *in, const int64_t vdims[3], const float *k_dev, const int kd[3], float *out,
               hipStream_t st)
{
    if (vdims[0] < 1 || vdims[1] < 1 || vdims[2] < 1) {
        set_error("convn: volume dimensions must be positive");
        return SID_ERR_DIMS;
    }
    if (kd[0] < 1 || kd[1] < 1 || kd[2] < 1 || kd[0] > kConvnMaxXY || kd[1] > kConvnMaxXY || kd[2] > kConvnMaxZ) {
        set_error("convn: kernel %d x %d x %d outside 1..%d x 1..%d x 1..%d", kd[0], kd[1], kd[2], kConvnMaxXY,
                  kConvnMaxXY, kConvnMaxZ);
        return SID_ERR_DIMS;
    }
    if (vdims[0] > (1 << 24) || vdims[1] > (1 << 24) || vdims[2] > 65535) {
        set_error("convn: volume too large");
        return SID_ERR_DIMS;
    }
    const size_t lds = sizeof(float) * ((size_t)(kCvTx + kd[0] - 1) * (kCvTy + kd[1] - 1) + (size_t)kd[0] * kd[1]);
    const dim3 grid((unsigned)((vdims[0] + kCvTx - 1) / kCvTx), (unsigned)((vdims[1] + kCvTy - 1) / kCvTy),
                    (unsigned)vdims[2]);
    hipLaunchKernelGGL(k_convn_same, grid, dim3(256), lds, st, in, (int)vdims[0], (int)vdims[1], (int)vdims[2], k_dev,
                       kd[0], kd[1], kd[2], out);
    SID_TRY_HIP(hipGetLastError());
    return SID_OK;
}

}  // namespace sid
