// Direct light-field projector on MFMA: the lenslet-sublattice PSF
// convolution of external/LFrecon_util/forwardProjectGPU.m:16-28 and
// backwardProjectGPU_new.m:16-26 written as implicit GEMMs.
//
// Forward.  With p = n*P + (u,v) (sensor pixel in lenslet P) and
// q = n*Q + (a,b) (voxel), the reference sum
//     O[p] = sum_c sum_q V[q,c] * H_{a,b,c}[p - q + off]
// becomes, for every lenslet tap D = P - Q,
//     O[P,uv] += sum_{c,ab} V[P - D, ab, c] * W_c[D][ab][uv],
//     W_c[D][ab][uv] = H_{a,b,c}[n*D + (u - a, v - b) + off].
// For a fixed plane c and tap row dx the sum over dy and ab is one GEMM
// whose A operand is Toeplitz along y: the K index runs over the input
// lenslets Qy = Py - dy in ascending order and ab inside each, so for output
// lenslet P the A row is a contiguous run of the padded volume row Px - dx.
// M = output lenslets, N = uv (padded to 16*NT), K = (#dy taps) * S.
//
// Backward (the adjoint, backwardProjectGPU_new.m:19-23):
//     B[Q,ab,c] = sum_D sum_uv O[Q + D, uv] * W_c[D][ab][uv]
// with the same structure: A = padded image rows Qx + dx, K = (dy, uv)
// ascending, N = ab.
//
// Every (plane, dx) pair is a segment; a block computes one M tile over a
// contiguous range of segments (split-K) into its own partial-sum slab; a
// reduce kernel sums the slabs in a fixed order (deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "sid_internal.h"

namespace sid {

// K chunk of the staged FP32 B operand (elements of K per LDS stage).
constexpr int kKcF32 = 32;

// ---------------------------------------------------------------------------
// Support scan: smallest index box holding every nonzero of a plane.
// ---------------------------------------------------------------------------
__global__ void k_support_scan(const float *__restrict__ H, int64_t M, int64_t n, int64_t nz,
                               int *__restrict__ box /* nz x 4: ilo, ihi, jlo, jhi */)
{
    const int64_t per_plane = M * M * n * n;
    const int c = blockIdx.y;
    int ilo = INT32_MAX, ihi = -1, jlo = INT32_MAX, jhi = -1;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < per_plane;
         e += (int64_t)gridDim.x * blockDim.x) {
        float v = H[(int64_t)c * per_plane + e];
        if (v != 0.0f) {
            int i = (int)(e % M), j = (int)((e / M) % M);
            ilo = min(ilo, i); ihi = max(ihi, i); jlo = min(jlo, j); jhi = max(jhi, j);
        }
    }
    // wave reduction, then one atomic per wave
    for (int o = 32; o > 0; o >>= 1) {
        ilo = min(ilo, __shfl_xor(ilo, o)); ihi = max(ihi, __shfl_xor(ihi, o));
        jlo = min(jlo, __shfl_xor(jlo, o)); jhi = max(jhi, __shfl_xor(jhi, o));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&box[c * 4 + 0], ilo); atomicMax(&box[c * 4 + 1], ihi);
        atomicMin(&box[c * 4 + 2], jlo); atomicMax(&box[c * 4 + 3], jhi);
    }
}

// ---------------------------------------------------------------------------
// Weight re-layout: H (MATLAB layout) -> per (plane, dx) GEMM B operands,
// stored chunk-transposed: [K / KC][NP][KC] so one LDS stage is one
// contiguous copy with K innermost.
// ---------------------------------------------------------------------------
// Entry idx of the weight block of plane c in direction dir: [tap row][K / KC][NP][KC],
// K = taps * Sk (tap ky holds rows ky * Sk .. ky * Sk + n^2 - 1, the rest zero).
__device__ __forceinline__ float w_value(const float *__restrict__ H, int64_t idx, int64_t M, int n, int c, int dir,
                                         int lo, int hi, int off, int t_lo, int ntap, int64_t K, int Sk, int NP,
                                         int KC)
{
    const int64_t per_tap = K * NP;
    const int nn2 = n * n;
    const int t_hi = t_lo + ntap - 1;
    const int tx = (int)(idx / per_tap);
    const int64_t rem = idx - (int64_t)tx * per_tap;
    const int64_t chunk = rem / ((int64_t)NP * KC);
    const int64_t r2 = rem - chunk * NP * KC;
    const int col = (int)(r2 / KC);
    const int64_t k = chunk * KC + (r2 - (int64_t)col * KC);
    const int dx = t_lo + tx;
    const int ky = (int)(k / Sk), kk = (int)(k - (int64_t)ky * Sk);
    if (ky < ntap && kk < nn2 && col < nn2) {
        int dy, ab, uv;
        if (dir == 0) { dy = t_hi - ky; ab = kk; uv = col; }
        else          { dy = t_lo + ky; uv = kk; ab = col; }
        const int a = ab % n, b = ab / n, u = uv % n, w = uv / n;
        const int i = n * dx + u - a + off, j = n * dy + w - b + off;
        if (i >= lo && i <= hi && j >= lo && j <= hi)
            return H[(int64_t)i + M * ((int64_t)j + M * ((int64_t)a + n * ((int64_t)b + (int64_t)n * c)))];
    }
    return 0.0f;
}

template <int PART>  // 0: fp32, 1: bf16 hi, 2: bf16 lo, 3: fp64
__global__ void k_build_w(const float *__restrict__ H, void *__restrict__ Wout, int64_t M, int n,
                          int c, int dir, int lo, int hi, int off, int t_lo, int ntap, int64_t K,
                          int S, int NP, int KC)
{
    const int64_t total = (int64_t)ntap * K * NP;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const float v = w_value(H, idx, M, n, c, dir, lo, hi, off, t_lo, ntap, K, S, NP, KC);
        if (PART == 0) {
            ((float *)Wout)[idx] = v;
        } else if (PART == 3) {
            ((double *)Wout)[idx] = (double)v;
        } else {
            u16 h = f2bf(v);
            ((u16 *)Wout)[idx] = PART == 1 ? h : f2bf(v - bf2f(h));
        }
    }
}

// Zero tiles of a weight layout: one wave per tile of E consecutive entries
// (BF16 layout: E = 512, 16 columns of a 32-row chunk; FP32 layout: E = NP * 32,
// a whole 32-row chunk); flag[t] = 1 when any entry is nonzero.  Both BF16 parts
// of an entry are zero exactly when its FP32 value is.
__global__ void k_tile_flags(const float *__restrict__ H, uint8_t *__restrict__ flag, int64_t ntiles, int E, int64_t M,
                             int n, int c, int dir, int lo, int hi, int off, int t_lo, int ntap, int64_t K, int Sk,
                             int NP)
{
    const int lane = threadIdx.x & 63;
    const int64_t t = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (t >= ntiles) return;   // whole waves
    bool nz = false;
    for (int e = lane; e < E && !nz; e += 64)
        nz = w_value(H, t * E + e, M, n, c, dir, lo, hi, off, t_lo, ntap, K, Sk, NP, 32) != 0.0f;
    const unsigned long long b = __ballot(nz);
    if (lane == 0) flag[t] = b != 0 ? 1 : 0;
}

// ---------------------------------------------------------------------------
// The FP32 projector GEMM.  256 threads = 4 waves; each wave owns RT*16 rows of
// the M tile and all NT*16 columns.  B chunks [NP][32] are staged in LDS and
// shared by the 4 waves; A is read straight from global memory (L2-resident
// rows, each row a contiguous Toeplitz window).  Chunks whose weights are all
// zero (marked in the segment's zero-tile table) are skipped.
//
// v_mfma_f32_16x16x4_f32: lane (l16, g) = (lane & 15, lane >> 4) feeds
// A[row l16][k = 8g + s] and B[k = 8g + s][col l16] at step s = 0..7 of a
// 32-wide K chunk, so a lane's 8 A values are contiguous.  C layout: lane
// holds C[4g + i][l16], i = 0..3.
// ---------------------------------------------------------------------------
template <int NT>
struct F32Shape {
    static constexpr int NP = NT * 16;
    static constexpr int KC = kKcF32;
    static constexpr int LDK = KC + 4;  // floats per staged column
};

// One work item; the barrier before each stage also orders it after the
// previous item's last LDS reads.
template <int NT, int RT>
__device__ __forceinline__ void f32_item(const Item it, float (&Bs)[F32Shape<NT>::NP * F32Shape<NT>::LDK],
                                         const Seg *__restrict__ segs, const uint32_t *__restrict__ tmap, int nkc,
                                         const float *__restrict__ A0, const float *__restrict__ W0,
                                         float *__restrict__ part, int nrows, int Ly, int YP, int S,
                                         int64_t slab_stride)
{
    constexpr int NP = F32Shape<NT>::NP;
    constexpr int MT = 4 * RT * 16;
    constexpr int KC = F32Shape<NT>::KC;
    constexpr int LDK = F32Shape<NT>::LDK;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, g = lane >> 4;

    int64_t rowoff[RT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
        int r = it.mtile * MT + wave * RT * 16 + rt * 16 + l16;
        r = r < nrows ? r : nrows - 1;
        const int Px = r / Ly, Py = r - Px * Ly;
        rowoff[rt] = ((int64_t)Px * YP + Py) * S;
    }
    f4 acc[RT][NT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) acc[rt][ct] = f4{0.f, 0.f, 0.f, 0.f};

    for (int s = it.seg0; s < it.seg1; ++s) {
        const Seg sg = segs[s];
        const float *A = A0 + sg.a_off;
        const float *W = W0 + sg.w_off;
        const uint32_t *live = tmap + sg.tmap + nkc;   // nonzero 32-row chunks, one bit each
        for (int k0 = 0; k0 < sg.K; k0 += KC) {
            const int j = k0 / KC;
            if (!((live[j >> 5] >> (j & 31)) & 1u)) continue;   // all-zero weights (uniform)
            float4 a0[RT], a1[RT];
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) {
                const float *pa = A + rowoff[rt] + k0 + 8 * g;
                a0[rt] = *(const float4 *)pa;
                a1[rt] = *(const float4 *)(pa + 4);
            }
            __syncthreads();
            const float *wc = W + (int64_t)k0 * NP;  // chunk: [NP][KC]
            for (int e = threadIdx.x; e < NP * KC / 4; e += NTHREADS) {
                const int col = e / (KC / 4), kq = e - col * (KC / 4);
                *(float4 *)(Bs + col * LDK + 4 * kq) = *(const float4 *)(wc + 4 * e);
            }
            __syncthreads();
#pragma unroll
            for (int ct = 0; ct < NT; ++ct) {
                const float *pb = Bs + (ct * 16 + l16) * LDK + 8 * g;
                const float4 b0 = *(const float4 *)pb, b1 = *(const float4 *)(pb + 4);
#pragma unroll
                for (int rt = 0; rt < RT; ++rt) {
                    f4 c = acc[rt][ct];
                    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[rt].x, b0.x, c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[rt].y, b0.y, c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[rt].z, b0.z, c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[rt].w, b0.w, c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[rt].x, b1.x, c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[rt].y, b1.y, c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[rt].z, b1.z, c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[rt].w, b1.w, c, 0, 0, 0);
                    acc[rt][ct] = c;
                }
            }
        }
    }

    float *out = part + (int64_t)it.slab * slab_stride;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
        const int rbase = it.mtile * MT + wave * RT * 16 + rt * 16 + 4 * g;
#pragma unroll
        for (int ct = 0; ct < NT; ++ct)
#pragma unroll
            for (int i = 0; i < 4; ++i) out[(int64_t)(rbase + i) * NP + ct * 16 + l16] = acc[rt][ct][i];
    }
}

// Blocks take items blockIdx.x, blockIdx.x + gridDim.x, ...: one item per
// block with the full grid, or a capped grid that leaves CUs to a kernel
// running beside it on another stream.
template <int NT, int RT>
__global__ __launch_bounds__(NTHREADS, 2) void k_proj_gemm(const Item *__restrict__ items, int nitems,
                                                        const Seg *__restrict__ segs,
                                                        const uint32_t *__restrict__ tmap, int nkc,
                                                        const float *__restrict__ A0, const float *__restrict__ W0,
                                                        float *__restrict__ part, int nrows, int Ly, int YP, int S,
                                                        int64_t slab_stride)
{
    __shared__ float Bs[F32Shape<NT>::NP * F32Shape<NT>::LDK];
    for (int i = blockIdx.x; i < nitems; i += gridDim.x) {
        const Item it = items[i];
        if (it.mtile >= 0)
            f32_item<NT, RT>(it, Bs, segs, tmap, nkc, A0, W0, part, nrows, Ly, YP, S, slab_stride);
    }
}


// ---------------------------------------------------------------------------
// BF16 projector GEMM with register-sliding A (MODE 1: BF16X2, MODE 2: BF16X3).
//
// Within a segment the A row of output lenslet r at tap ky is the window of
// lenslet (r + ky) of the padded volume row, so lane r's A fragment at tap
// ky + 1 is lane r + 1's at tap ky whenever lenslets r and r + 1 are
// neighbours in the same row (always except at the end of a volume row).  The
// loop runs over K chunks of 32 (kc, outer) and taps (ky, inner): after each
// tap the A registers move one lane down (DPP row_shl:1) and only lane 15 and
// lanes at a row end load their next fragment from memory, a tap ahead of its
// use.  Each tap is padded to KT = 32 * ceil(S / 32) in the weights (rows
// beyond S are zero), so the fragments past S read the next lenslet's values
// and multiply zeros.  B chunks [NP][32] are double buffered in LDS through
// registers, one barrier per step.  Chunk kc runs over the taps of the
// segment's tap table only (the hull of its nonzero weight tiles): at the edge
// taps of a plane whole chunks of weights are zero, and skipping them changes
// no sum of finite values.
// ---------------------------------------------------------------------------
typedef int i32x4 __attribute__((ext_vector_type(4)));

// Lane i of each 16-lane row takes lane i + 1's value (DPP row_shl:1); lane 15 keeps its own.
__device__ __forceinline__ bf8 dpp_row_shl1(bf8 v)
{
    i32x4 x = __builtin_bit_cast(i32x4, v);
    x.x = __builtin_amdgcn_update_dpp(x.x, x.x, 0x101, 0xf, 0xf, false);
    x.y = __builtin_amdgcn_update_dpp(x.y, x.y, 0x101, 0xf, 0xf, false);
    x.z = __builtin_amdgcn_update_dpp(x.z, x.z, 0x101, 0xf, 0xf, false);
    x.w = __builtin_amdgcn_update_dpp(x.w, x.w, 0x101, 0xf, 0xf, false);
    return __builtin_bit_cast(bf8, x);
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// B chunk [NP][KW] (16-byte vectors e = col * KW / 8 + kq) of the hi (and lo) weights into registers.
template <int NP, int KW, int NB, int VPT>
__device__ __forceinline__ void sl_fetch_b(u32x4 (&pre)[NB][VPT], const u16 *__restrict__ Wh,
                                           const u16 *__restrict__ Wl, int64_t o)
{
    constexpr int NV = NP * KW / 8;
#pragma unroll
    for (int v = 0; v < VPT; ++v) {
        // threads past the chunk load its last vector again (the stash skips
        // them): an unconditional load keeps pre out of the loop's phi copies
        const int e = std::min(v * NTHREADS + (int)threadIdx.x, NV - 1);
        pre[0][v] = *(const u32x4 *)(Wh + o + 8 * e);
        if constexpr (NB == 2) pre[1][v] = *(const u32x4 *)(Wl + o + 8 * e);
    }
}

// ... and from registers into an LDS buffer [NB][NP * LDK] (column pitch LDK).
template <int NP, int KW, int LDK, int NB, int VPT>
__device__ __forceinline__ void sl_stash_b(const u32x4 (&pre)[NB][VPT], u16 *bs)
{
    constexpr int NV = NP * KW / 8;
#pragma unroll
    for (int v = 0; v < VPT; ++v) {
        // threads past the chunk hold its last vector (sl_fetch_b) and store it
        // again at the same place: the same bytes, no branch
        const int e = std::min(v * NTHREADS + (int)threadIdx.x, NV - 1);
        const int col = e / (KW / 8), kq = e - col * (KW / 8);
        *(u32x4 *)(bs + col * LDK + 8 * kq) = pre[0][v];
        if constexpr (NB == 2) *(u32x4 *)(bs + NP * LDK + col * LDK + 8 * kq) = pre[1][v];
    }
}

template <int NT, int MODE>
struct SlShape {
    static constexpr int NP = NT * 16;
    static constexpr int KW = 32;            // K per step
    static constexpr int LDK = KW + 8;       // staged column pitch (BF16 elements)
    static constexpr int NB = MODE == 2 ? 2 : 1;
};

// Buffer resource over the A operand from `base` (gfx9 word 3): a lane whose
// offset is kNoLoad reads zeros without a memory access.
constexpr uint32_t kNoLoad = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t a_rsrc(const u16 *base)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, 0x7fffffff, 0x00020000);
}

// Decoded step-table entry (uniform): A offset, B offset, and whether the step
// continues the previous step's chunk at the next tap.
struct SlStep {
    int64_t a, w;
    bool slide;
};

// Entry whose 32-bit word l % 4 lane l holds.
__device__ __forceinline__ SlStep sl_decode(int v)
{
    const uint32_t a0 = __builtin_amdgcn_readlane(v, 0), a1 = __builtin_amdgcn_readlane(v, 1);
    const uint32_t w0 = __builtin_amdgcn_readlane(v, 2), w1 = __builtin_amdgcn_readlane(v, 3);
    const uint64_t wv = (uint64_t)w1 << 32 | w0;
    return SlStep{(int64_t)((uint64_t)a1 << 32 | a0), (int64_t)(wv >> 1), (wv & 1u) != 0};
}

// One work item (an M tile over a run of segments) of the sliding GEMM.  Bs is
// the block's LDS; the barrier before the first stash also orders this item's
// staging after the previous item's last reads.  The item's steps (one tap of
// one K chunk each) come from the run's step table (A offset; B offset << 1 |
// 1 when the step continues the previous step's chunk at the next tap).  All
// steps of all chunks and segments form one software pipeline; while step t
// runs its MFMAs,
//   - the B chunk of step t + 1 waits in LDS and that of step t + 2 is in flight,
//   - the A fragments of step t + 1 are in flight: every lane loads at the first
//     tap of a chunk; at the next taps only lane 15 and the lanes at a volume
//     row end load (the others pass an offset past the buffer resource's range,
//     which reads zeros without a memory access) and the rest take the step's
//     fragments slid one lane down (DPP row_shl:1), chosen per lane at the top
//     of step t + 1,
//   - the table entry of step t + 3 is in flight.
// Every load is unconditional in the data flow and each loaded register is
// consumed before it is loaded again, so no copy (and no wait for a load just
// issued) sits in the loop.  Lane l reads 32-bit word l % 4 of a table entry
// (a vector load, read out with readlane): a scalar-memory wait in the loop
// would also drain the outstanding LDS reads.
template <int NT, int RT, int MODE>
__device__ __forceinline__ void sl_item(
    const Item it, u16 (&Bs)[2][SlShape<NT, MODE>::NB][SlShape<NT, MODE>::NP * SlShape<NT, MODE>::LDK],
    const int64_t *__restrict__ steps, const int32_t *__restrict__ run0, const u16 *__restrict__ A0,
    const u16 *__restrict__ A1, const u16 *__restrict__ W0, const u16 *__restrict__ W1, float *__restrict__ part,
    int nrows, int Ly, int YP, int S, int64_t slab_stride)
{
    using SH = SlShape<NT, MODE>;
    constexpr int NP = SH::NP;
    constexpr int MT = 4 * RT * 16;
    constexpr int KW = SH::KW;
    constexpr int LDK = SH::LDK;
    constexpr int NB = SH::NB;
    constexpr int NV = NP * KW / 8;      // 16-byte vectors per staged part
    constexpr int VPT = (NV + NTHREADS - 1) / NTHREADS;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, g = lane >> 4;

    uint32_t aoff[RT];   // byte offset of the lane's A window from a step's A offset
    bool fresh[RT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
        const int r0 = it.mtile * MT + wave * RT * 16 + rt * 16 + l16;
        const int r = r0 < nrows ? r0 : nrows - 1;
        const int rn = r0 + 1 < nrows ? r0 + 1 : nrows - 1;
        const int Px = r / Ly, Py = r - Px * Ly;
        const int Pxn = rn / Ly, Pyn = rn - Pxn * Ly;
        const int64_t ro = ((int64_t)Px * YP + Py) * S;
        const int64_t next = ((int64_t)Pxn * YP + Pyn) * S;
        aoff[rt] = (uint32_t)((ro + 8 * g) * 2);
        fresh[rt] = l16 == 15 || next != ro + S;
    }
    f4 acc[RT][NT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) acc[rt][ct] = f4{0.f, 0.f, 0.f, 0.f};

    const int step0 = run0[it.slab], nst = run0[it.slab + 1] - step0;
    if (nst > 0) {
        const int *ent = (const int *)(steps + 2 * (int64_t)step0) + (lane & 3);
        // A fragments of step e (all lanes, or only the fresh ones when it slides)
        auto load_a = [&](const SlStep &e, bf8 (&h)[RT], bf8 (&l)[RT]) {
            const __amdgpu_buffer_rsrc_t rh = a_rsrc(A0 + e.a), rl = a_rsrc(A1 + e.a);
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) {
                const uint32_t o = (!e.slide || fresh[rt]) ? aoff[rt] : kNoLoad;
                h[rt] = __builtin_bit_cast(bf8, __builtin_amdgcn_raw_buffer_load_b128(rh, o, 0, 0));
                l[rt] = __builtin_bit_cast(bf8, __builtin_amdgcn_raw_buffer_load_b128(rl, o, 0, 0));
            }
        };
        // The loop body is one basic block (entries past the run are clamped to
        // its last step, whose loads are then repeated and unused, and the last
        // step stages an unused B chunk), so the compiler's wait counts stay
        // exact across iterations and no load sinks below the MFMAs.
        const int last = nst - 1;
        u32x4 pre[NB][VPT];
        bf8 ldh[RT], ldl[RT], slh[RT], sll[RT];
        const int v1 = ent[4 * min(1, last)], v2 = ent[4 * min(2, last)];
        const SlStep e0 = sl_decode(ent[0]);
        load_a(e0, ldh, ldl);
        sl_fetch_b<NP, KW, NB, VPT>(pre, W0, W1, e0.w);
        bool slide_t = false;   // the first step of a run starts a chunk
        SlStep e_1 = sl_decode(v1), e_2 = sl_decode(v2);   // entries of steps t + 1, t + 2
        int vn = ent[4 * min(3, last)];                   // entry word of step t + 3, in flight
        __syncthreads();   // previous item's buffers consumed
        sl_stash_b<NP, KW, LDK, NB, VPT>(pre, &Bs[0][0][0]);
        sl_fetch_b<NP, KW, NB, VPT>(pre, W0, W1, e_1.w);
        __syncthreads();
        for (int t = 0, buf = 0; t < nst; ++t, buf ^= 1) {
            // A fragments of step t: loaded, or the previous step's slid one lane down
            bf8 ah[RT], al[RT];
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) {
                const bool ld = !slide_t || fresh[rt];
                ah[rt] = ld ? ldh[rt] : slh[rt];
                al[rt] = ld ? ldl[rt] : sll[rt];
                slh[rt] = dpp_row_shl1(ah[rt]);
                sll[rt] = dpp_row_shl1(al[rt]);
            }
            load_a(e_1, ldh, ldl);   // A of step t + 1
            {
                // B fragments through a ring of kBR register slots (NB fragments each):
                // fragment ct + kBR is read right after the MFMAs of fragment ct are
                // issued, so each LDS read has kBR - 1 fragments' MFMAs to land under
                // (the scheduling groups keep that order).
                constexpr int kBR = 3;
                constexpr int kMf = (MODE == 2 ? 3 : 2) * RT;   // MFMAs per fragment
                const u16 *bb = &Bs[buf][0][l16 * LDK + 8 * g];
                bf8 ring[kBR][NB];
#pragma unroll
                for (int i = 0; i < kBR; ++i)
#pragma unroll
                    for (int h = 0; h < NB; ++h) ring[i][h] = *(const bf8 *)(bb + h * NP * LDK + i * 16 * LDK);
                __builtin_amdgcn_sched_group_barrier(0x020, 2 * RT, 0);     // the A loads of step t + 1 first
                __builtin_amdgcn_sched_group_barrier(0x100, kBR * NB, 0);   // the first reads
#pragma unroll
                for (int ct = 0; ct < NT; ++ct) {
                    const bf8 bh = ring[ct % kBR][0], bl = ring[ct % kBR][NB - 1];
#pragma unroll
                    for (int rt = 0; rt < RT; ++rt) {
                        f4 c = acc[rt][ct];
                        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rt], bh, c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[rt], bh, c, 0, 0, 0);
                        if constexpr (MODE == 2) c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rt], bl, c, 0, 0, 0);
                        acc[rt][ct] = c;
                    }
                    __builtin_amdgcn_sched_group_barrier(0x008, kMf, 0);   // this fragment's MFMAs
                    if (ct + kBR < NT) {
#pragma unroll
                        for (int h = 0; h < NB; ++h)
                            ring[ct % kBR][h] = *(const bf8 *)(bb + h * NP * LDK + (ct + kBR) * 16 * LDK);
                        __builtin_amdgcn_sched_group_barrier(0x100, NB, 0);   // then the next reads
                    }
                }
            }
            sl_stash_b<NP, KW, LDK, NB, VPT>(pre, &Bs[buf ^ 1][0][0]);   // B of step t + 1
            slide_t = e_1.slide;
            sl_fetch_b<NP, KW, NB, VPT>(pre, W0, W1, e_2.w);               // B of step t + 2
            e_1 = e_2;
            e_2 = sl_decode(vn);
            vn = ent[4 * min(t + 4, last)];
            __syncthreads();
        }
    }

    float *out = part + (int64_t)it.slab * slab_stride;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
        const int rbase = it.mtile * MT + wave * RT * 16 + rt * 16 + 4 * g;
#pragma unroll
        for (int ct = 0; ct < NT; ++ct)
#pragma unroll
            for (int i = 0; i < 4; ++i) out[(int64_t)(rbase + i) * NP + ct * 16 + l16] = acc[rt][ct][i];
    }
}

// Blocks take items blockIdx.x, blockIdx.x + gridDim.x, ...: one item per
// block with the full grid, or a capped grid that leaves CUs to a kernel
// running beside it on another stream.
template <int NT, int RT, int MODE>
__global__ __launch_bounds__(NTHREADS, (MODE == 2 || RT > 2) ? 1 : 2) void k_proj_gemm_sl(
    const Item *__restrict__ items, int nitems, const int64_t *__restrict__ steps, const int32_t *__restrict__ run0,
    const u16 *__restrict__ A0, const u16 *__restrict__ A1, const u16 *__restrict__ W0, const u16 *__restrict__ W1,
    float *__restrict__ part, int nrows, int Ly, int YP, int S, int64_t slab_stride)
{
    using SH = SlShape<NT, MODE>;
    __shared__ u16 Bs[2][SH::NB][SH::NP * SH::LDK];
    for (int i = blockIdx.x; i < nitems; i += gridDim.x) {
        const Item it = items[i];
        if (it.mtile >= 0)
            sl_item<NT, RT, MODE>(it, Bs, steps, run0, A0, A1, W0, W1, part, nrows, Ly, YP, S, slab_stride);
    }
}

// FP64 variant (reference path for error studies): v_mfma_f64_16x16x4_f64
// with the same K permutation as MODE 0 (lane group g supplies k = 8g + s at
// step s of a 32-wide chunk).  Its output layout differs from the FP32 one:
// item i of lane (l16, g) is C[4i + g][l16].  It skips the all-zero chunks as
// the FP32 GEMM does (same layout, same chunk masks).
template <int NT, int RT>
__global__ __launch_bounds__(NTHREADS) void k_proj_gemm_f64(
    const Item *__restrict__ items, const Seg *__restrict__ segs, const uint32_t *__restrict__ tmap, int nkc,
    const double *__restrict__ A0, const double *__restrict__ W0, double *__restrict__ part, int nrows, int Ly,
    int YP, int S, int64_t slab_stride)
{
    constexpr int NP = NT * 16;
    constexpr int MT = 4 * RT * 16;
    constexpr int KC = 32;
    constexpr int LDK = KC + 2;
    __shared__ double Bs[NP * LDK];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, g = lane >> 4;
    const Item it = items[blockIdx.x];
    if (it.mtile < 0) return;  // padding of the per-XCD queues (whole block)
    int64_t rowoff[RT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
        int r = it.mtile * MT + wave * RT * 16 + rt * 16 + l16;
        r = r < nrows ? r : nrows - 1;
        const int Px = r / Ly, Py = r - Px * Ly;
        rowoff[rt] = ((int64_t)Px * YP + Py) * S;
    }
    d4 acc[RT][NT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) acc[rt][ct] = d4{0., 0., 0., 0.};
    for (int s = it.seg0; s < it.seg1; ++s) {
        const Seg sg = segs[s];
        const double *A = A0 + sg.a_off;
        const double *W = W0 + sg.w_off;
        const uint32_t *live = tmap + sg.tmap + nkc;   // nonzero 32-row chunks, one bit each
        for (int k0 = 0; k0 < sg.K; k0 += KC) {
            const int j = k0 / KC;
            if (!((live[j >> 5] >> (j & 31)) & 1u)) continue;   // all-zero weights (uniform)
            double2 a[RT][4];
#pragma unroll
            for (int rt = 0; rt < RT; ++rt)
#pragma unroll
                for (int q = 0; q < 4; ++q) a[rt][q] = *(const double2 *)(A + rowoff[rt] + k0 + 8 * g + 2 * q);
            __syncthreads();
            const double *wc = W + (int64_t)k0 * NP;
            for (int e = threadIdx.x; e < NP * KC / 2; e += NTHREADS) {
                const int col = e / (KC / 2), kq = e - col * (KC / 2);
                *(double2 *)(Bs + col * LDK + 2 * kq) = *(const double2 *)(wc + 2 * e);
            }
            __syncthreads();
#pragma unroll
            for (int ct = 0; ct < NT; ++ct) {
                double2 b[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) b[q] = *(const double2 *)(Bs + (ct * 16 + l16) * LDK + 8 * g + 2 * q);
#pragma unroll
                for (int rt = 0; rt < RT; ++rt) {
                    d4 c = acc[rt][ct];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt][q].x, b[q].x, c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt][q].y, b[q].y, c, 0, 0, 0);
                    }
                    acc[rt][ct] = c;
                }
            }
        }
    }
    // v_mfma_f64_16x16x4_f64 output layout: lane (l16, g), item i holds C[4i + g][l16]
    double *out = part + (int64_t)it.slab * slab_stride;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
        const int rbase = it.mtile * MT + wave * RT * 16 + rt * 16 + g;
#pragma unroll
        for (int ct = 0; ct < NT; ++ct)
#pragma unroll
            for (int i = 0; i < 4; ++i) out[(int64_t)(rbase + 4 * i) * NP + ct * 16 + l16] = acc[rt][ct][i];
    }
}

// ---------------------------------------------------------------------------
// Reductions and layout conversions
// ---------------------------------------------------------------------------

__device__ __forceinline__ void store_shadow(u16 *hi, u16 *lo, int64_t o, float v)
{
    const u16 h = f2bf(v);
    hi[o] = h;
    lo[o] = f2bf(v - bf2f(h));
}

// Sum the forward slabs; write the padded lenslet-major image (T = float:
// plus BF16 hi/lo shadows when hi != nullptr) and, if O_col != nullptr, the
// MATLAB column-major image.  Pixels beyond X, Y (the last partial lenslet)
// are 0.
// Non-finite input (stamp `check` in nf[0]): every sensor pixel is NaN, as one
// NaN or Inf voxel makes forwardProjectGPU's FFT convolutions NaN everywhere
// (conv2FFT_.m:14).  mark != 0: a non-finite output pixel stamps nf[1] (the
// input of the backward that follows).
template <typename T>
__global__ void k_fwd_reduce(const T *__restrict__ part, int nslabs, int64_t slab_stride, int NP, int nrows,
                             int n, int Ly, int YP, int T_, int S, int X, int Y, T *__restrict__ img,
                             u16 *__restrict__ hi, u16 *__restrict__ lo, float *__restrict__ O_col,
                             const float2 *__restrict__ fft_img, int fft_ny, const float *__restrict__ tout,
                             const float *__restrict__ tout2, unsigned long long *__restrict__ nf,
                             unsigned long long check, unsigned long long mark)
{
    const int nn2 = n * n;
    const int64_t total = (int64_t)nrows * nn2;
    const bool poison = check != 0 && nf[0] == check;
    bool bad = false;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(e / nn2), uv = (int)(e - (int64_t)r * nn2);
        T s = 0;
        const T *p = part + (int64_t)r * NP + uv;
        for (int k = 0; k < nslabs; ++k) s += p[(int64_t)k * slab_stride];
        const int Px = r / Ly, Py = r - Px * Ly;
        const int x = Px * n + uv % n, y = Py * n + uv / n;
        if (fft_img) s += (T)fft_img[(int64_t)x * fft_ny + y].x;   // planes on the FFT path
        if (tout) s += (T)tout[(int64_t)r * NP + uv];                 // ... in the tiled form
        if (tout2) s += (T)tout2[(int64_t)r * NP + uv];               // ... its second chain
        if (poison) s = (T)NAN;
        if (x >= X || y >= Y) s = 0;
        bad = bad || !isfinite(s);
        const int64_t o = ((int64_t)(Px + T_) * YP + (Py + T_)) * S + uv;
        img[o] = s;
        if (hi) store_shadow(hi, lo, o, (float)s);
        if (O_col && x < X && y < Y) O_col[(int64_t)x + (int64_t)X * y] = (float)s;
    }
    if (bad && mark) nf[1] = mark;
}

// Sum the backward slabs of each plane.  Without rl_update: write the MATLAB
// column-major volume B_col.  With rl_update (deconvRL.m:52,85-87):
//   r = Htf / D; r(isinf) = 0; X = X .* r; X(isnan) = 0
// on the padded volume (and its BF16 shadows), and copy X to B_col if given.
// Backward reduction: block (r, c) = lenslet row r of plane c, thread ab
// (n*n <= 256 phases, contiguous in the padded layout): sum of the row's
// split-K slabs (direct planes) or the FFT-path result, then either the
// column-major output or the ISRA update X <- X * Htf / B (deconvRL.m:51-57
// guards: Inf ratio -> 0, NaN product -> 0).
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_reduce(const T *__restrict__ part, const int *__restrict__ plane_slab0,
                                                    int64_t slab_stride, int NP, int n, int Ly, int64_t XP,
                                                    int64_t YP, int T_, int S, int X, int Y,
                                                    float *__restrict__ B_col, int rl_update,
                                                    const T *__restrict__ htf, T *__restrict__ vol,
                                                    u16 *__restrict__ hi, u16 *__restrict__ lo,
                                                    const int *__restrict__ plane_fft,
                                                    const float *__restrict__ dfft,
                                                    unsigned long long *__restrict__ nf,
                                                    unsigned long long check, unsigned long long mark,
                                                    const int *__restrict__ planes)
{
    const int ab = threadIdx.x;
    if (ab >= n * n) return;
    const int r = blockIdx.x, c = planes ? planes[blockIdx.y] : (int)blockIdx.y;
    const int Px = r / Ly, Py = r - Px * Ly;
    const int64_t o = (((int64_t)c * XP + Px + T_) * YP + (Py + T_)) * S + ab;
    T d = 0;
    if (plane_fft && plane_fft[c]) {
        d = (T)dfft[o];   // plane on the FFT path
    } else {
        const T *p = part + (int64_t)r * NP + ab;
        const int s0 = plane_slab0[c], s1 = plane_slab0[c + 1];
        for (int k = s0; k < s1; ++k) d += p[(int64_t)k * slab_stride];
    }
    if (check != 0 && nf[1] == check) d = (T)NAN;   // non-finite image: NaN everywhere
    const int x = Px * n + ab % n, y = Py * n + ab / n;
    const bool real = x < X && y < Y;
    const int64_t col = (int64_t)x + (int64_t)X * ((int64_t)y + (int64_t)Y * c);
    if (!rl_update) {
        if (real) B_col[col] = (float)d;
        return;
    }
    T v = 0;
    if (real) {
        T ratio = htf[o] / d;
        if (isinf(ratio)) ratio = 0;
        v = vol[o] * ratio;
        if (isnan(v)) v = 0;
    }
    if (mark && !isfinite(v)) nf[0] = mark;   // an Inf iterate poisons the next forward
    vol[o] = v;
    if (hi) store_shadow(hi, lo, o, (float)v);
    if (B_col && real) B_col[col] = (float)v;
}

// MATLAB column-major volume -> padded lenslet-major volume (+ BF16 shadows),
// through LDS: a block takes one lenslet row Py of one plane and kVinPx
// lenslets along x, reads the n sensor rows of that strip whole (n * kVinPx
// contiguous values each) and writes each lenslet record (n * n contiguous
// values) whole; a thread per element wrote n-value fragments of records.
constexpr int kVinPx = 16;
template <typename T>
__global__ __launch_bounds__(256) void k_vol_in_t(const float *__restrict__ V, int X, int Y, int n, int Lx,
                                                  int64_t XP, int64_t YP, int T_, int S, T *__restrict__ dst,
                                                  u16 *__restrict__ hi, u16 *__restrict__ lo,
                                                  unsigned long long *__restrict__ nf, unsigned long long mark)
{
    __shared__ float tile[16 * 16 * kVinPx];   // n <= 16 rows of n * kVinPx values
    const int Py = blockIdx.x, Px0 = blockIdx.y * kVinPx, c = blockIdx.z;
    const int W = n * kVinPx, x0 = Px0 * n, y0 = Py * n;
    const float *src = V + (int64_t)c * X * Y;
    bool bad = false;
    for (int e = threadIdx.x; e < n * W; e += blockDim.x) {
        const int r = e / W, col = e - r * W;
        const int x = x0 + col, y = y0 + r;
        const float v = (x < X && y < Y) ? src[(int64_t)y * X + x] : 0.0f;
        bad = bad || !isfinite(v);
        tile[e] = v;
    }
    if (bad && mark) nf[0] = mark;
    __syncthreads();
    const int n2 = n * n;
    for (int e = threadIdx.x; e < kVinPx * n2; e += blockDim.x) {
        const int k = e / n2, ab = e - k * n2;
        const int a = ab % n, b = ab / n;
        const int Px = Px0 + k;
        if (Px >= Lx || Px * n + a >= X || y0 + b >= Y) continue;
        const float v = tile[b * W + k * n + a];
        const int64_t o = (((int64_t)c * XP + Px + T_) * YP + (Py + T_)) * S + ab;
        if (dst) dst[o] = (T)v;
        if (hi) store_shadow(hi, lo, o, v);
    }
}

template <typename T>
__global__ void k_vol_in(const float *__restrict__ V, int X, int Y, int nz, int n, int64_t XP, int64_t YP, int T_,
                         int S, T *__restrict__ dst, u16 *__restrict__ hi, u16 *__restrict__ lo,
                         unsigned long long *__restrict__ nf_slot, unsigned long long mark)
{
    const int64_t total = (int64_t)X * Y * nz;
    bool bad = false;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(e % X);
        const int64_t t = e / X;
        const int y = (int)(t % Y), c = (int)(t / Y);
        const int Px = x / n, Py = y / n, ab = (x - Px * n) + n * (y - Py * n);
        const int64_t o = (((int64_t)c * XP + Px + T_) * YP + (Py + T_)) * S + ab;
        const float v = V[e];
        bad = bad || !isfinite(v);
        if (dst) dst[o] = (T)v;
        if (hi) store_shadow(hi, lo, o, v);
    }
    if (bad && mark) *nf_slot = mark;
}

// Padded lenslet-major volume -> MATLAB column-major volume.
template <typename T>
__global__ void k_vol_out(const T *__restrict__ src, int X, int Y, int nz, int n, int64_t XP, int64_t YP, int T_,
                          int S, float *__restrict__ V)
{
    const int64_t total = (int64_t)X * Y * nz;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(e % X);
        const int64_t t = e / X;
        const int y = (int)(t % Y), c = (int)(t / Y);
        const int Px = x / n, Py = y / n, ab = (x - Px * n) + n * (y - Py * n);
        V[e] = (float)src[(((int64_t)c * XP + Px + T_) * YP + (Py + T_)) * S + ab];
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

static inline int floordiv(int a, int b) { return (a >= 0) ? a / b : -((-a + b - 1) / b); }
static inline int64_t roundup(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

static int grid_for(int64_t total, int threads = 256, int cap = 65536)
{
    int64_t g = (total + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

int support_scan(sid_psf *p)
{
    const int64_t nz = p->nzH;
    std::vector<int> box(nz * 4);
    for (int64_t c = 0; c < nz; ++c) {
        box[c * 4 + 0] = INT32_MAX; box[c * 4 + 1] = -1;
        box[c * 4 + 2] = INT32_MAX; box[c * 4 + 3] = -1;
    }
    int *d_box = nullptr;
    SID_TRY_HIP(hipMalloc(&d_box, box.size() * sizeof(int)));
    SID_TRY_HIP(hipMemcpyAsync(d_box, box.data(), box.size() * sizeof(int), hipMemcpyHostToDevice, p->stream));
    const int64_t per_plane = p->M * p->M * p->n * p->n;
    dim3 grid((unsigned)std::min<int64_t>((per_plane + 255) / 256, 1024), (unsigned)nz);
    hipLaunchKernelGGL(k_support_scan, grid, dim3(256), 0, p->stream, p->H, p->M, p->n, nz, d_box);
    SID_TRY_HIP(hipGetLastError());
    SID_TRY_HIP(hipMemcpyAsync(box.data(), d_box, box.size() * sizeof(int), hipMemcpyDeviceToHost, p->stream));
    SID_TRY_HIP(hipStreamSynchronize(p->stream));
    SID_TRY_HIP(hipFree(d_box));
    for (int64_t c = 0; c < nz; ++c) {
        Plane &pl = p->planes[c];
        if (box[c * 4 + 1] < 0) {
            pl.empty = true;
            pl.lo = 0;
            pl.hi = -1;
        } else {
            pl.lo = std::min(box[c * 4 + 0], box[c * 4 + 2]);
            pl.hi = std::max(box[c * 4 + 1], box[c * 4 + 3]);
            pl.empty = false;
        }
    }
    return SID_OK;
}

// Tap range for offsets d = i - off, i in [lo, hi]: n*D + (u - a) = d with
// |u - a| <= n - 1.
static void tap_range(int lo, int hi, int off, int n, int *t_lo, int *t_hi)
{
    *t_lo = -floordiv(-(lo - off - (n - 1)), n);  // ceil
    *t_hi = floordiv(hi - off + (n - 1), n);
}

void free_weights(sid_psf *p)
{
    for (void **q : {(void **)&p->wf, (void **)&p->wb, (void **)&p->wf_hi, (void **)&p->wb_hi, (void **)&p->wf_lo,
                     (void **)&p->wb_lo, (void **)&p->wf_d, (void **)&p->wb_d}) {
        if (*q) (void)hipFree(*q);
        *q = nullptr;
    }
}

// Zero-tile flags, from H: p->tflag_f / tflag_b for the 32 x 16 tiles of the
// BF16 layouts, p->cflag_f / cflag_b for the 32-row chunks of the FP32 layouts.
static int scan_tiles(sid_psf *p)
{
    const int n = (int)p->n;
    const int64_t EC = p->NP * 32;
    p->tflag_f.assign((size_t)(p->wf2_elems / 512), 0);
    p->tflag_b.assign((size_t)(p->wb2_elems / 512), 0);
    p->cflag_f.assign((size_t)(p->wf_elems / EC), 0);
    p->cflag_b.assign((size_t)(p->wb_elems / EC), 0);
    uint8_t *d_flag = nullptr;
    const int64_t nmax = std::max<int64_t>({(int64_t)p->tflag_f.size(), (int64_t)p->tflag_b.size(),
                                            (int64_t)p->cflag_f.size(), (int64_t)p->cflag_b.size(), (int64_t)1});
    SID_TRY_HIP(hipMalloc(&d_flag, nmax));
    int rc = SID_OK;
    for (int lay = 0; lay < 4 && rc == SID_OK; ++lay) {   // BF16 fwd, BF16 bwd, FP32 fwd, FP32 bwd
        const int dir = lay & 1;
        const bool bf = lay < 2;
        const int64_t E = bf ? 512 : EC;
        for (int64_t c = 0; c < p->nzH; ++c) {
            const Plane &pl = p->planes[c];
            if (pl.empty) continue;
            const int t_lo = dir == 0 ? pl.tf_lo : pl.tb_lo;
            const int ntap = (dir == 0 ? pl.tf_hi : pl.tb_hi) - t_lo + 1;
            const int64_t K = bf ? (dir == 0 ? pl.Kf2 : pl.Kb2) : (dir == 0 ? pl.Kf : pl.Kb);
            const int64_t w0 = bf ? (dir == 0 ? pl.wf2_off : pl.wb2_off) : (dir == 0 ? pl.wf_off : pl.wb_off);
            const int64_t nt = (int64_t)ntap * K * p->NP / E;
            hipLaunchKernelGGL(k_tile_flags, dim3((unsigned)((nt + 3) / 4)), dim3(256), 0, p->stream, p->H,
                               d_flag + w0 / E, nt, (int)E, p->M, n, (int)c, dir, pl.lo, pl.hi,
                               dir == 0 ? pl.off_f : pl.off_b, t_lo, ntap, K, bf ? (int)p->KT : (int)p->S, (int)p->NP);
            if (hipGetLastError() != hipSuccess) { rc = SID_ERR_HIP; break; }
        }
        std::vector<uint8_t> &f = lay == 0 ? p->tflag_f : lay == 1 ? p->tflag_b : lay == 2 ? p->cflag_f : p->cflag_b;
        if (rc == SID_OK && !f.empty() &&
            (hipMemcpyAsync(f.data(), d_flag, f.size(), hipMemcpyDeviceToHost, p->stream) != hipSuccess ||
             hipStreamSynchronize(p->stream) != hipSuccess))
            rc = SID_ERR_HIP;
    }
    (void)hipFree(d_flag);
    if (rc != SID_OK) set_error("zero-tile scan of the weights failed");
    return rc;
}

int build_weights(sid_psf *p, int which)
{
    const int n = (int)p->n;
    // geometry (once)
    if (p->wf_elems == 0) {
        int64_t offf = 0, offb = 0, off2f = 0, off2b = 0;
        p->T = 0;
        p->KT = roundup(p->S, 32);
        for (auto &pl : p->planes) {
            if (pl.empty) {
                pl.tf_lo = pl.tb_lo = 0;
                pl.tf_hi = pl.tb_hi = -1;
                pl.Kf = pl.Kb = pl.Kf2 = pl.Kb2 = 0;
                pl.wf_off = offf;
                pl.wb_off = offb;
                pl.wf2_off = off2f;
                pl.wb2_off = off2b;
                continue;
            }
            tap_range(pl.lo, pl.hi, pl.off_f, n, &pl.tf_lo, &pl.tf_hi);
            tap_range(pl.lo, pl.hi, pl.off_b, n, &pl.tb_lo, &pl.tb_hi);
            const int ntf = pl.tf_hi - pl.tf_lo + 1, ntb = pl.tb_hi - pl.tb_lo + 1;
            pl.Kf = roundup((int64_t)ntf * p->S, KPAD);
            pl.Kb = roundup((int64_t)ntb * p->S, KPAD);
            pl.wf_off = offf;
            pl.wb_off = offb;
            offf += (int64_t)ntf * pl.Kf * p->NP;
            offb += (int64_t)ntb * pl.Kb * p->NP;
            pl.Kf2 = (int64_t)ntf * p->KT;
            pl.Kb2 = (int64_t)ntb * p->KT;
            pl.wf2_off = off2f;
            pl.wb2_off = off2b;
            off2f += (int64_t)ntf * pl.Kf2 * p->NP;
            off2b += (int64_t)ntb * pl.Kb2 * p->NP;
            p->T = std::max<int64_t>(p->T, std::max(std::max(std::abs(pl.tf_lo), std::abs(pl.tf_hi)),
                                                    std::max(std::abs(pl.tb_lo), std::abs(pl.tb_hi))));
        }
        p->wf_elems = offf;
        p->wb_elems = offb;
        p->wf2_elems = off2f;
        p->wb2_elems = off2b;
    }
    const size_t esz = which == 0 ? 4 : which == 3 ? 8 : 2;
    void **dstf = which == 0 ? (void **)&p->wf : which == 1 ? (void **)&p->wf_hi
                : which == 2 ? (void **)&p->wf_lo : (void **)&p->wf_d;
    void **dstb = which == 0 ? (void **)&p->wb : which == 1 ? (void **)&p->wb_hi
                : which == 2 ? (void **)&p->wb_lo : (void **)&p->wb_d;
    if (*dstf) return SID_OK;
    // FP32 / FP64: [K / 32][NP][32] with K = taps * S (padded); BF16: the same with each tap padded to KT
    const bool bf = which == 1 || which == 2;
    const int KC = bf ? 32 : kKcF32;
    const int64_t ef = bf ? p->wf2_elems : p->wf_elems, eb = bf ? p->wb2_elems : p->wb_elems;
    const int Sk = bf ? (int)p->KT : (int)p->S;
    SID_TRY_HIP(hipMalloc(dstf, std::max<int64_t>(ef, 1) * esz));
    SID_TRY_HIP(hipMalloc(dstb, std::max<int64_t>(eb, 1) * esz));
    for (int64_t c = 0; c < p->nzH; ++c) {
        const Plane &pl = p->planes[c];
        if (pl.empty) continue;
        for (int dir = 0; dir < 2; ++dir) {
            const int t_lo = dir == 0 ? pl.tf_lo : pl.tb_lo;
            const int ntap = (dir == 0 ? pl.tf_hi : pl.tb_hi) - t_lo + 1;
            const int64_t K = bf ? (dir == 0 ? pl.Kf2 : pl.Kb2) : (dir == 0 ? pl.Kf : pl.Kb);
            const int off = dir == 0 ? pl.off_f : pl.off_b;
            const int64_t woff = bf ? (dir == 0 ? pl.wf2_off : pl.wb2_off) : (dir == 0 ? pl.wf_off : pl.wb_off);
            char *base = (char *)(dir == 0 ? *dstf : *dstb) + woff * esz;
            const int64_t total = (int64_t)ntap * K * p->NP;
            const int grid = grid_for(total, 256, 16384);
            if (which == 0)
                hipLaunchKernelGGL(k_build_w<0>, dim3(grid), dim3(256), 0, p->stream, p->H, (void *)base, p->M, n,
                                   (int)c, dir, pl.lo, pl.hi, off, t_lo, ntap, K, Sk, (int)p->NP, KC);
            else if (which == 1)
                hipLaunchKernelGGL(k_build_w<1>, dim3(grid), dim3(256), 0, p->stream, p->H, (void *)base, p->M, n,
                                   (int)c, dir, pl.lo, pl.hi, off, t_lo, ntap, K, Sk, (int)p->NP, KC);
            else if (which == 2)
                hipLaunchKernelGGL(k_build_w<2>, dim3(grid), dim3(256), 0, p->stream, p->H, (void *)base, p->M, n,
                                   (int)c, dir, pl.lo, pl.hi, off, t_lo, ntap, K, Sk, (int)p->NP, KC);
            else
                hipLaunchKernelGGL(k_build_w<3>, dim3(grid), dim3(256), 0, p->stream, p->H, (void *)base, p->M, n,
                                   (int)c, dir, pl.lo, pl.hi, off, t_lo, ntap, K, Sk, (int)p->NP, KC);
            SID_TRY_HIP(hipGetLastError());
        }
    }
    SID_TRY_HIP(hipStreamSynchronize(p->stream));
    if (which == 0) SID_TRY(scan_tiles(p));
    return SID_OK;
}

// 16-row MFMA tiles per wave (the block's M tile is 4 * 16 * RT rows).  RT = 4
// for BF16X2 (occupancy 1, 120 MFMAs per barrier) measured 23 % slower alone
// and 16 % slower beside the FFT path than RT = 2.
static int row_tiles_for(int /*precision*/) { return 2; }

// Split a list of segments into `nsplit` contiguous groups of near-equal total cost.
static std::vector<std::pair<int, int>> split_segments(const std::vector<int64_t> &cost, int s0, int s1, int nsplit)
{
    std::vector<std::pair<int, int>> out;
    int64_t total = 0;
    for (int s = s0; s < s1; ++s) total += cost[s];
    if (s1 <= s0) return out;
    nsplit = std::max(1, std::min(nsplit, s1 - s0));
    int start = s0;
    int64_t acc = 0;
    for (int s = s0; s < s1; ++s) {
        acc += cost[s];
        const int g = (int)out.size();
        const int64_t target = total * (g + 1) / nsplit;
        const int remaining_groups = nsplit - g - 1;
        const int remaining_segs = s1 - s - 1;
        if ((acc >= target && remaining_segs >= remaining_groups) || remaining_segs == remaining_groups) {
            if (g < nsplit - 1) {
                out.push_back({start, s + 1});
                start = s + 1;
            }
        }
    }
    if (start < s1) out.push_back({start, s1});
    return out;
}

// gfx950 has 8 XCDs, each with its own L2; the dispatcher deals workgroup b to
// XCD b % 8, and a grid-stride block of a grid that is a multiple of 8 keeps
// taking items on its XCD.  So position p of an item list runs on XCD p % 8.
constexpr int kXcds = 8;

// Item list for the XCDs: each group (the M tiles of one run of segments,
// which all read the same weight chunks) goes whole to one XCD, costliest
// first onto the least-loaded XCD, so one L2 fetches a group's weights once;
// the per-XCD queues (costliest groups first) are interleaved, the shorter
// ones padded with empty items (mtile < 0) that the kernels skip.
static std::vector<Item> xcd_order(const std::vector<std::vector<Item>> &groups, const std::vector<int64_t> &cost)
{
    std::vector<int> idx(groups.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
        return cost[a] * (int64_t)groups[a].size() > cost[b] * (int64_t)groups[b].size();
    });
    std::vector<std::vector<Item>> q(kXcds);
    std::vector<int64_t> load(kXcds, 0);
    for (int g : idx) {
        const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        q[x].insert(q[x].end(), groups[g].begin(), groups[g].end());
        load[x] += cost[g] * (int64_t)groups[g].size();
    }
    size_t len = 0;
    for (auto &v : q) len = std::max(len, v.size());
    std::vector<Item> out(len * kXcds, Item{0, 0, -1, 0});
    for (int x = 0; x < kXcds; ++x)
        for (size_t j = 0; j < q[x].size(); ++j) out[j * kXcds + x] = q[x][j];
    return out;
}

// Zero-tile table of a segment, appended to tab (returns its offset).
// BF16 GEMM: word kc (< nkc) holds ky0 | ky1 << 16, the hull [ky0, ky1) of the
// taps whose 32-row chunk kc has a nonzero weight tile (0 when none has); the
// GEMM runs chunk kc over these taps only.  Skipping single 16-column tiles
// inside a tap as well (a uniform branch per fragment) measured 10 % slower on
// 1005^2 x 51 planes: the branches break up the interleaving of LDS reads and
// MFMAs.  FP32 GEMM: words nkc, nkc + 1, ... hold one bit per 32-row chunk of
// the segment's K (chunk j: bit j % 32 of word nkc + j / 32), set when the
// chunk has a nonzero weight; the GEMM skips the others.
static int32_t seg_table(std::vector<uint32_t> &tab, const std::vector<uint8_t> &tflag,
                         const std::vector<uint8_t> &cflag, const Seg &sg, int ntap, int nkc, int NT, int64_t NP)
{
    const size_t at = tab.size();
    const int nch = sg.K / 32;
    tab.resize(at + nkc + (nch + 31) / 32, 0);
    const int64_t t0 = sg.w_off2 / 512;
    for (int kc = 0; kc < nkc; ++kc) {
        int ky0 = ntap, ky1 = 0;
        for (int ky = 0; ky < ntap; ++ky) {
            bool any = false;
            for (int ct = 0; ct < NT && !any; ++ct) {
                const int64_t t = t0 + (int64_t)(ky * nkc + kc) * NT + ct;
                any = t >= (int64_t)tflag.size() || tflag[t];   // unknown: computed
            }
            if (any) {
                ky0 = std::min(ky0, ky);
                ky1 = ky + 1;
            }
        }
        tab[at + kc] = ky0 < ky1 ? (uint32_t)ky0 | (uint32_t)ky1 << 16 : 0u;
    }
    const int64_t c0 = sg.w_off / (NP * 32);
    for (int j = 0; j < nch; ++j) {
        const int64_t t = c0 + j;
        if (t >= (int64_t)cflag.size() || cflag[t]) tab[at + nkc + j / 32] |= 1u << (j % 32);
    }
    return (int32_t)at;
}

// Step table of the BF16 GEMM for the run of segments [s0, s1): one entry per
// (segment, chunk kc, tap ky of the chunk's hull), in the order the GEMM runs
// them, appended to g.steps.  Entry = (A offset of the tap's window, B offset
// of chunk (ky, kc) << 1 | 1 when ky continues the previous entry's chunk).
// Within each plane of the run the chunk kc is the outer loop and the tap
// rows (segments) the inner one: the A windows of tap rows dx and dx + 1 are
// the same lenslet rows shifted by one, so the M tiles of a group that run
// together on one XCD re-read a chunk's 32 elements of the plane's lenslet
// records from L2 instead of fetching the plane once per tap row.
static void append_steps(GemmPlan &g, int s0, int s1, int64_t S, int64_t NP, int nkc)
{
    for (int p0 = s0; p0 < s1;) {
        int p1 = p0 + 1;
        while (p1 < s1 && g.segs[p1].plane == g.segs[p0].plane) ++p1;
        for (int kc = 0; kc < nkc; ++kc)
            for (int s = p0; s < p1; ++s) {
                const Seg &sg = g.segs[s];
                const uint32_t h = g.tmap[sg.tmap + kc];
                const int ky0 = (int)(h & 0xffffu), ky1 = (int)(h >> 16);
                for (int ky = ky0; ky < ky1; ++ky) {
                    g.steps.push_back(sg.a_off + (int64_t)ky * S + (int64_t)kc * 32);
                    g.steps.push_back((sg.w_off2 + ((int64_t)ky * nkc + kc) * NP * 32) << 1 | (ky > ky0 ? 1 : 0));
                }
            }
        p0 = p1;
    }
    g.run0.push_back((int32_t)(g.steps.size() / 2));
}

static int upload_plan(GemmPlan &g)
{
    if (g.d_segs) { (void)hipFree(g.d_segs); g.d_segs = nullptr; }
    if (g.d_items) { (void)hipFree(g.d_items); g.d_items = nullptr; }
    if (g.d_plane_slab0) { (void)hipFree(g.d_plane_slab0); g.d_plane_slab0 = nullptr; }
    if (g.d_tmap) { (void)hipFree(g.d_tmap); g.d_tmap = nullptr; }
    if (g.d_steps) { (void)hipFree(g.d_steps); g.d_steps = nullptr; }
    if (g.d_run0) { (void)hipFree(g.d_run0); g.d_run0 = nullptr; }
    SID_TRY_HIP(hipMalloc(&g.d_steps, std::max<size_t>(g.steps.size(), 2) * sizeof(int64_t)));
    SID_TRY_HIP(hipMalloc(&g.d_run0, std::max<size_t>(g.run0.size(), 1) * sizeof(int32_t)));
    if (!g.steps.empty())
        SID_TRY_HIP(hipMemcpy(g.d_steps, g.steps.data(), g.steps.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    if (!g.run0.empty())
        SID_TRY_HIP(hipMemcpy(g.d_run0, g.run0.data(), g.run0.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    SID_TRY_HIP(hipMalloc(&g.d_tmap, std::max<size_t>(g.tmap.size(), 1) * sizeof(uint32_t)));
    if (!g.tmap.empty())
        SID_TRY_HIP(hipMemcpy(g.d_tmap, g.tmap.data(), g.tmap.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    SID_TRY_HIP(hipMalloc(&g.d_segs, std::max<size_t>(g.segs.size(), 1) * sizeof(Seg)));
    SID_TRY_HIP(hipMalloc(&g.d_items, std::max<size_t>(g.items.size(), 1) * sizeof(Item)));
    if (!g.segs.empty())
        SID_TRY_HIP(hipMemcpy(g.d_segs, g.segs.data(), g.segs.size() * sizeof(Seg), hipMemcpyHostToDevice));
    if (!g.items.empty())
        SID_TRY_HIP(hipMemcpy(g.d_items, g.items.data(), g.items.size() * sizeof(Item), hipMemcpyHostToDevice));
    if (!g.plane_slab0.empty()) {
        SID_TRY_HIP(hipMalloc(&g.d_plane_slab0, g.plane_slab0.size() * sizeof(int)));
        SID_TRY_HIP(hipMemcpy(g.d_plane_slab0, g.plane_slab0.data(), g.plane_slab0.size() * sizeof(int),
                              hipMemcpyHostToDevice));
    }
    return SID_OK;
}

void free_plan_buffers(SensorPlan &sp)
{
    for (void *q : {(void *)sp.vol, (void *)sp.img, (void *)sp.htf, (void *)sp.part, (void *)sp.vol_hi,
                    (void *)sp.vol_lo, (void *)sp.img_hi, (void *)sp.img_lo, (void *)sp.vol_d, (void *)sp.img_d,
                    (void *)sp.htf_d, (void *)sp.part_d})
        if (q) (void)hipFree(q);
    sp.vol = sp.img = sp.htf = sp.part = nullptr;
    sp.vol_hi = sp.vol_lo = sp.img_hi = sp.img_lo = nullptr;
    sp.vol_d = sp.img_d = sp.htf_d = sp.part_d = nullptr;
    sp.prec_built = -1;
}

int ensure_plan(sid_psf *p, int64_t X, int64_t Y, int64_t nz)
{
    SensorPlan &sp = p->plan;
    const int RT = row_tiles_for(p->precision);
    const int64_t mt = 4 * RT * 16;
    // planes on the FFT path (the FP64 reference path is direct only)
    std::vector<char> use(p->nzH, 0);
    if (p->precision != SID_PREC_FP64) {
        for (int64_t c = 0; c < p->nzH; ++c) {
            if (!fft_eligible(p, c)) continue;
            if (p->method == SID_METHOD_FFT) use[c] = 1;
            else if (p->method != SID_METHOD_AUTO) continue;
            else if ((size_t)c < p->pin.size() && p->pin[c]) use[c] = p->pin[c] == SID_METHOD_FFT;
            else if ((size_t)c < p->auto_fft.size()) use[c] = p->auto_fft[c];
        }
    }
    // the FFT state stores FP16 spectra in the BF16X2 mode and FP32 otherwise
    const bool fft_ok = !sp.fft.active || sp.fft.half == (p->precision == SID_PREC_BF16X2);
    // the GEMM work split follows the cost of the precision's GEMM
    const int cost_bf16 = p->precision == SID_PREC_BF16X2 || p->precision == SID_PREC_BF16X3;
    if (sp.X == X && sp.Y == Y && sp.nz == nz && sp.mt == mt && sp.vol && use == sp.fft_use && fft_ok &&
        sp.cost_bf16 == cost_bf16)
        return ensure_precision_buffers(p);
    free_plan_buffers(sp);
    sp.cost_bf16 = cost_bf16;
    fft_state_free(sp.fft);
    sp.fft_use = use;
    const int64_t n = p->n, S = p->S, NP = p->NP, T = p->T;
    sp.X = X; sp.Y = Y; sp.nz = nz;
    sp.Lx = (X + n - 1) / n; sp.Ly = (Y + n - 1) / n;
    sp.XP = sp.Lx + 2 * T; sp.YP = sp.Ly + 2 * T + 1;
    sp.nrows = sp.Lx * sp.Ly;
    if (sp.XP * sp.YP * S * 2 >= (1LL << 31) - 64) {   // BF16 GEMM lane offsets are 32-bit byte offsets
        set_error("sensor %lld x %lld too large for the projector's 32-bit plane offsets", (long long)X, (long long)Y);
        return SID_ERR_DIMS;
    }
    sp.mt = mt;
    sp.ntiles = (sp.nrows + mt - 1) / mt;
    sp.mpad = sp.ntiles * mt;
    int dev_cus = 256;
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, p->device) == hipSuccess) dev_cus = prop.multiProcessorCount;
    }
    // Work items per CU the segment split aims for.  Alone on the device the
    // GEMM gets 16 (at 1005^2 x 51 planes, alternating runs: 8 -> 16 took FP32
    // from 9.23 to 9.63 and BF16X2 from 53.2 to 54.7 RL iter/s, the 16-frame
    // gradient unchanged; 4 had left a third wave of blocks 11-25 % full);
    // beside the FFT path 6 (BF16X2, alternating runs on one device: 4 -> 113.9 /
    // 115.1, 6 -> 118.9 / 119.5, 3 -> 118.3 / 117.3 RL iter/s; 8 measured 11 %
    // slower than 4 before the tap hulls).  SID_SPLIT_PER_CU overrides both.
    const bool any_fft = std::find(use.begin(), use.end(), 1) != use.end();
    int per_cu = any_fft ? 6 : 16;
    if (const char *e = std::getenv("SID_SPLIT_PER_CU")) per_cu = std::max(1, std::atoi(e));
    const int target_blocks = per_cu * dev_cus;
    // Segment cost: the K rows the GEMM runs (the nonzero chunks for FP32, the tap hulls for BF16).
    const int nkc = (int)(p->KT / 32);
    auto seg_cost = [&](const GemmPlan &g, const Seg &sg) -> int64_t {
        int64_t k = 0;
        if (!cost_bf16) {
            for (int w = 0; w < (sg.K / 32 + 31) / 32; ++w) k += 32 * (int64_t)__builtin_popcount(g.tmap[sg.tmap + nkc + w]);
            return k;
        }
        for (int kc = 0; kc < nkc; ++kc) {
            const uint32_t h = g.tmap[sg.tmap + kc];
            k += 32 * (int64_t)((int)(h >> 16) - (int)(h & 0xffffu));
        }
        return k;
    };

    // ---- forward: segments over planes 0..nz-1 and their dx rows
    GemmPlan &F = sp.fwd;
    F.segs.clear(); F.items.clear(); F.tmap.clear(); F.cost.clear(); F.steps.clear(); F.run0.assign(1, 0); F.macs = 0;
    for (int64_t c = 0; c < nz; ++c) {
        const Plane &pl = p->planes[c];
        if (pl.empty || use[c]) continue;
        const int ntap = pl.tf_hi - pl.tf_lo + 1;
        for (int tx = 0; tx < ntap; ++tx) {
            const int dx = pl.tf_lo + tx;
            Seg sg;
            sg.a_off = ((c * sp.XP + (T - dx)) * sp.YP + (T - pl.tf_hi)) * S;
            sg.w_off = pl.wf_off + (int64_t)tx * pl.Kf * NP;
            sg.K = (int32_t)pl.Kf;
            sg.plane = (int32_t)c;
            sg.w_off2 = pl.wf2_off + (int64_t)tx * pl.Kf2 * NP;
            sg.K2 = (int32_t)pl.Kf2;
            sg.tmap = seg_table(F.tmap, p->tflag_f, p->cflag_f, sg, ntap, nkc, (int)(NP / 16), NP);
            F.segs.push_back(sg);
            F.cost.push_back(seg_cost(F, sg));
        }
    }
    {
        // a multiple of the XCD count of equal-cost groups balances the XCDs
        int nsplit = (int)std::max<int64_t>(1, (target_blocks + sp.ntiles - 1) / sp.ntiles);
        nsplit = (nsplit + kXcds - 1) / kXcds * kXcds;
        auto groups = split_segments(F.cost, 0, (int)F.segs.size(), nsplit);
        F.nslabs = (int)groups.size();
        std::vector<std::vector<Item>> gi_items(groups.size());
        std::vector<int64_t> gi_cost(groups.size(), 0);
        for (int gi = 0; gi < (int)groups.size(); ++gi) {
            for (int t = 0; t < sp.ntiles; ++t) gi_items[gi].push_back(Item{groups[gi].first, groups[gi].second, t, gi});
            for (int s = groups[gi].first; s < groups[gi].second; ++s) gi_cost[gi] += F.cost[s];
            append_steps(F, groups[gi].first, groups[gi].second, S, NP, nkc);   // run of slab gi
        }
        F.items = xcd_order(gi_items, gi_cost);
        // MFMA multiply-adds the GEMM runs: the executed K rows (tap hulls / live chunks) x N x padded M
        for (size_t i = 0; i < F.segs.size(); ++i) F.macs += F.cost[i] * NP * sp.mpad;
    }

    // ---- backward: every plane of H, split per plane in proportion to cost
    GemmPlan &B = sp.bwd;
    B.segs.clear(); B.items.clear(); B.tmap.clear(); B.cost.clear(); B.plane_slab0.assign(p->nzH + 1, 0); B.macs = 0;
    B.steps.clear(); B.run0.assign(1, 0);
    std::vector<int> pseg0(p->nzH + 1, 0);
    int64_t btotal = 0;
    for (int64_t c = 0; c < p->nzH; ++c) {
        pseg0[c] = (int)B.segs.size();
        const Plane &pl = p->planes[c];
        if (pl.empty || use[c]) continue;
        const int ntap = pl.tb_hi - pl.tb_lo + 1;
        for (int tx = 0; tx < ntap; ++tx) {
            const int dx = pl.tb_lo + tx;
            Seg sg;
            sg.a_off = ((T + dx) * sp.YP + (T + pl.tb_lo)) * S;
            sg.w_off = pl.wb_off + (int64_t)tx * pl.Kb * NP;
            sg.K = (int32_t)pl.Kb;
            sg.plane = (int32_t)c;
            sg.w_off2 = pl.wb2_off + (int64_t)tx * pl.Kb2 * NP;
            sg.K2 = (int32_t)pl.Kb2;
            sg.tmap = seg_table(B.tmap, p->tflag_b, p->cflag_b, sg, ntap, nkc, (int)(NP / 16), NP);
            B.segs.push_back(sg);
            B.cost.push_back(seg_cost(B, sg));
            btotal += B.cost.back();
        }
    }
    pseg0[p->nzH] = (int)B.segs.size();
    {
        const int64_t per_split = std::max<int64_t>(1, btotal * sp.ntiles / target_blocks);
        int slab = 0;
        std::vector<std::vector<Item>> gr_items;
        std::vector<int64_t> gr_cost;
        for (int64_t c = 0; c < p->nzH; ++c) {
            B.plane_slab0[c] = slab;
            int64_t cost = 0;
            for (int s = pseg0[c]; s < pseg0[c + 1]; ++s) cost += B.cost[s];
            if (cost == 0) continue;
            const int nsplit = (int)std::max<int64_t>(1, (cost + per_split - 1) / per_split);
            auto groups = split_segments(B.cost, pseg0[c], pseg0[c + 1], nsplit);
            for (auto &gr : groups) {
                int64_t gc = 0;
                for (int s = gr.first; s < gr.second; ++s) gc += B.cost[s];
                gr_items.emplace_back();
                for (int t = 0; t < sp.ntiles; ++t) gr_items.back().push_back(Item{gr.first, gr.second, t, slab});
                gr_cost.push_back(gc);
                append_steps(B, gr.first, gr.second, S, NP, nkc);   // run of this slab
                ++slab;
            }
        }
        B.plane_slab0[p->nzH] = slab;
        B.nslabs = slab;
        B.items = xcd_order(gr_items, gr_cost);
        for (size_t i = 0; i < B.segs.size(); ++i) B.macs += B.cost[i] * NP * sp.mpad;
    }
    SID_TRY(upload_plan(F));
    SID_TRY(upload_plan(B));

    // ---- buffers
    sp.vol_elems = p->nzH * sp.XP * sp.YP * S;
    sp.img_elems = sp.XP * sp.YP * S;
    sp.part_elems = (int64_t)std::max(std::max(F.nslabs, B.nslabs), 1) * sp.mpad * NP;
    SID_TRY_HIP(hipMalloc(&sp.vol, sp.vol_elems * sizeof(float)));
    SID_TRY_HIP(hipMalloc(&sp.htf, sp.vol_elems * sizeof(float)));
    SID_TRY_HIP(hipMalloc(&sp.img, sp.img_elems * sizeof(float)));
    SID_TRY_HIP(hipMalloc(&sp.part, sp.part_elems * sizeof(float)));
    SID_TRY_HIP(hipMemsetAsync(sp.vol, 0, sp.vol_elems * sizeof(float), p->stream));
    SID_TRY_HIP(hipMemsetAsync(sp.htf, 0, sp.vol_elems * sizeof(float), p->stream));
    SID_TRY_HIP(hipMemsetAsync(sp.img, 0, sp.img_elems * sizeof(float), p->stream));
    SID_TRY_HIP(hipStreamSynchronize(p->stream));
    bool any = false;
    for (char u : use) any = any || u;
    if (any) SID_TRY(fft_state_build(p, use, p->stream));
    return ensure_precision_buffers(p);
}

int ensure_precision_buffers(sid_psf *p)
{
    SensorPlan &sp = p->plan;
    if (p->precision == SID_PREC_FP32) return SID_OK;
    if (p->precision == SID_PREC_FP64) {
        SID_TRY(build_weights(p, 3));
        if (!sp.vol_d) {
            SID_TRY_HIP(hipMalloc(&sp.vol_d, sp.vol_elems * sizeof(double)));
            SID_TRY_HIP(hipMalloc(&sp.htf_d, sp.vol_elems * sizeof(double)));
            SID_TRY_HIP(hipMalloc(&sp.img_d, sp.img_elems * sizeof(double)));
            SID_TRY_HIP(hipMalloc(&sp.part_d, sp.part_elems * sizeof(double)));
            SID_TRY_HIP(hipMemsetAsync(sp.vol_d, 0, sp.vol_elems * sizeof(double), p->stream));
            SID_TRY_HIP(hipMemsetAsync(sp.htf_d, 0, sp.vol_elems * sizeof(double), p->stream));
            SID_TRY_HIP(hipMemsetAsync(sp.img_d, 0, sp.img_elems * sizeof(double), p->stream));
            SID_TRY_HIP(hipStreamSynchronize(p->stream));
        }
        return SID_OK;
    }
    SID_TRY(build_weights(p, 1));
    if (p->precision == SID_PREC_BF16X3) SID_TRY(build_weights(p, 2));
    if (!sp.vol_hi) {
        SID_TRY_HIP(hipMalloc(&sp.vol_hi, sp.vol_elems * sizeof(u16)));
        SID_TRY_HIP(hipMalloc(&sp.vol_lo, sp.vol_elems * sizeof(u16)));
        SID_TRY_HIP(hipMalloc(&sp.img_hi, sp.img_elems * sizeof(u16)));
        SID_TRY_HIP(hipMalloc(&sp.img_lo, sp.img_elems * sizeof(u16)));
        SID_TRY_HIP(hipMemsetAsync(sp.vol_hi, 0, sp.vol_elems * sizeof(u16), p->stream));
        SID_TRY_HIP(hipMemsetAsync(sp.vol_lo, 0, sp.vol_elems * sizeof(u16), p->stream));
        SID_TRY_HIP(hipMemsetAsync(sp.img_hi, 0, sp.img_elems * sizeof(u16), p->stream));
        SID_TRY_HIP(hipMemsetAsync(sp.img_lo, 0, sp.img_elems * sizeof(u16), p->stream));
        SID_TRY_HIP(hipStreamSynchronize(p->stream));
    }
    return SID_OK;
}

template <int NT>
static int launch_nt(sid_psf *p, const GemmPlan &g, int dir, hipStream_t st, int max_blocks)
{
    const SensorPlan &sp = p->plan;
    const int S = (int)p->S;
    const bool fwd = dir == 0;
    const int nitems = (int)g.items.size();
    const unsigned grid = (unsigned)(max_blocks > 0 ? std::min(nitems, max_blocks) : nitems);
    const int64_t slab = sp.mpad * (int64_t)(NT * 16);
    if (p->precision == SID_PREC_FP64) {
        hipLaunchKernelGGL((k_proj_gemm_f64<NT, 2>), dim3((unsigned)nitems), dim3(NTHREADS), 0, st, g.d_items,
                           g.d_segs, g.d_tmap, (int)(p->KT / 32), fwd ? sp.vol_d : sp.img_d, fwd ? p->wf_d : p->wb_d,
                           sp.part_d, (int)sp.nrows, (int)sp.Ly, (int)sp.YP, S, slab);
    } else if (p->precision == SID_PREC_FP32) {
        hipLaunchKernelGGL((k_proj_gemm<NT, 2>), dim3(grid), dim3(NTHREADS), 0, st, g.d_items, nitems, g.d_segs,
                           g.d_tmap, (int)(p->KT / 32), fwd ? sp.vol : sp.img, fwd ? p->wf : p->wb, sp.part,
                           (int)sp.nrows, (int)sp.Ly, (int)sp.YP, S, slab);
    } else {
        const u16 *A0 = fwd ? sp.vol_hi : sp.img_hi, *A1 = fwd ? sp.vol_lo : sp.img_lo;
        const u16 *W0 = fwd ? p->wf_hi : p->wb_hi;
        const u16 *W1 = p->precision == SID_PREC_BF16X3 ? (fwd ? p->wf_lo : p->wb_lo) : W0;
        if (p->precision == SID_PREC_BF16X2)
            hipLaunchKernelGGL((k_proj_gemm_sl<NT, 2, 1>), dim3(grid), dim3(NTHREADS), 0, st, g.d_items, nitems,
                               g.d_steps, g.d_run0, A0, A1, W0, W1, sp.part, (int)sp.nrows, (int)sp.Ly, (int)sp.YP, S,
                               slab);
        else
            hipLaunchKernelGGL((k_proj_gemm_sl<NT, 2, 2>), dim3(grid), dim3(NTHREADS), 0, st, g.d_items, nitems,
                               g.d_steps, g.d_run0, A0, A1, W0, W1, sp.part, (int)sp.nrows, (int)sp.Ly, (int)sp.YP, S,
                               slab);
    }
    SID_TRY_HIP(hipGetLastError());
    return SID_OK;
}

int launch_gemm(sid_psf *p, int dir, hipStream_t st, int max_blocks)
{
    const GemmPlan &g = dir == 0 ? p->plan.fwd : p->plan.bwd;
    if (g.items.empty()) return SID_OK;
    switch (p->NP / 16) {
        case 1: return launch_nt<1>(p, g, dir, st, max_blocks);
        case 2: return launch_nt<2>(p, g, dir, st, max_blocks);
        case 4: return launch_nt<4>(p, g, dir, st, max_blocks);
        case 8: return launch_nt<8>(p, g, dir, st, max_blocks);
        case 11: return launch_nt<11>(p, g, dir, st, max_blocks);
        case 15: return launch_nt<15>(p, g, dir, st, max_blocks);
        case 16: return launch_nt<16>(p, g, dir, st, max_blocks);
        default:
            set_error("unsupported padded lenslet size NP=%lld", (long long)p->NP);
            return SID_ERR_DIMS;
    }
}

// The BF16 hi/lo shadows of the volume and the image feed only the direct BF16
// GEMM: a plan with every nonempty plane on the FFT path writes none.
static bool bf16_shadows(const sid_psf *p)
{
    if (p->precision != SID_PREC_BF16X2 && p->precision != SID_PREC_BF16X3) return false;
    return !(p->plan.fft.active && p->plan.fft.all_fft);
}

int launch_fwd_reduce(sid_psf *p, float *O_col, hipStream_t st, unsigned long long check, unsigned long long mark)
{
    SensorPlan &sp = p->plan;
    const bool f64 = p->precision == SID_PREC_FP64;
    const bool sh = bf16_shadows(p);
    const int64_t total = sp.nrows * p->n * p->n;
    const int64_t stride = sp.mpad * p->NP;
    if (sp.fwd.nslabs == 0) {  // no contributing plane: the image is zero
        if (f64) SID_TRY_HIP(hipMemsetAsync(sp.part_d, 0, stride * sizeof(double), st));
        else SID_TRY_HIP(hipMemsetAsync(sp.part, 0, stride * sizeof(float), st));
    }
    const int ns = std::max(sp.fwd.nslabs, 1);
    if (f64)
        hipLaunchKernelGGL(k_fwd_reduce<double>, dim3(grid_for(total)), dim3(256), 0, st, sp.part_d, ns, stride,
                           (int)p->NP, (int)sp.nrows, (int)p->n, (int)sp.Ly, (int)sp.YP, (int)p->T, (int)p->S,
                           (int)sp.X, (int)sp.Y, sp.img_d, (u16 *)nullptr, (u16 *)nullptr, O_col,
                           (const float2 *)nullptr, 0, (const float *)nullptr, (const float *)nullptr, p->nf, check,
                           mark);
    else
        hipLaunchKernelGGL(k_fwd_reduce<float>, dim3(grid_for(total)), dim3(256), 0, st, sp.part, ns, stride,
                           (int)p->NP, (int)sp.nrows, (int)p->n, (int)sp.Ly, (int)sp.YP, (int)p->T, (int)p->S,
                           (int)sp.X, (int)sp.Y, sp.img, sh ? sp.img_hi : nullptr, sh ? sp.img_lo : nullptr, O_col,
                           sp.fft.active && !sp.fft.tiled ? (const float2 *)sp.fft.grid : nullptr, sp.fft.Ny,
                           sp.fft.active && sp.fft.tiled ? sp.fft.tout : nullptr,
                           sp.fft.active && sp.fft.tiled ? sp.fft.tout2 : nullptr, p->nf, check, mark);
    SID_TRY_HIP(hipGetLastError());
    return SID_OK;
}

// subset: 0 every plane of H, 1 the planes on the direct GEMM, 2 those on the
// FFT path (1 and 2 need an active FFT state; together they cover every plane).
int launch_bwd_reduce(sid_psf *p, float *B_col, int rl_update, hipStream_t st, unsigned long long check,
                      unsigned long long mark, int subset)
{
    SensorPlan &sp = p->plan;
    const bool f64 = p->precision == SID_PREC_FP64;
    const bool sh = bf16_shadows(p);
    const int64_t stride = sp.mpad * p->NP;
    const int *planes = nullptr;
    int64_t np = p->nzH;
    if (subset != 0) {
        if (!sp.fft.active || f64) {
            set_error("internal: a plane subset of the backward reduction needs the FFT path");
            return SID_ERR_ARG;
        }
        planes = sp.fft.d_plane_list + (subset == 2 ? sp.fft.n_list_direct : 0);
        np = subset == 2 ? sp.fft.n_list_fft : sp.fft.n_list_direct;
        if (np == 0) return SID_OK;
    }
    const dim3 grid((unsigned)sp.nrows, (unsigned)np);
    if (f64)
        hipLaunchKernelGGL(k_bwd_reduce<double>, grid, dim3(256), 0, st, sp.part_d, sp.bwd.d_plane_slab0, stride,
                           (int)p->NP, (int)p->n, (int)sp.Ly, sp.XP, sp.YP, (int)p->T, (int)p->S, (int)sp.X,
                           (int)sp.Y, B_col, rl_update, sp.htf_d, sp.vol_d, (u16 *)nullptr, (u16 *)nullptr,
                           (const int *)nullptr, (const float *)nullptr, p->nf, check, mark, (const int *)nullptr);
    else
        hipLaunchKernelGGL(k_bwd_reduce<float>, grid, dim3(256), 0, st, sp.part, sp.bwd.d_plane_slab0, stride,
                           (int)p->NP, (int)p->n, (int)sp.Ly, sp.XP, sp.YP, (int)p->T, (int)p->S, (int)sp.X,
                           (int)sp.Y, B_col, rl_update, sp.htf, sp.vol, sh ? sp.vol_hi : nullptr,
                           sh ? sp.vol_lo : nullptr, sp.fft.active ? sp.fft.d_plane_fft : nullptr,
                           sp.fft.active ? sp.fft.dfft : nullptr, p->nf, check, mark, planes);
    SID_TRY_HIP(hipGetLastError());
    return SID_OK;
}

// which: 0 = the iterate / forward input (with BF16 shadows), 1 = Htf.
int launch_vol_in(sid_psf *p, const float *V_col, int64_t nz, int which, hipStream_t st, unsigned long long mark)
{
    SensorPlan &sp = p->plan;
    const bool f64 = p->precision == SID_PREC_FP64;
    const bool sh = which == 0 && bf16_shadows(p);
    const int64_t total = sp.X * sp.Y * nz;
    if (total == 0) return SID_OK;
    const dim3 grid((unsigned)sp.Ly, (unsigned)((sp.Lx + kVinPx - 1) / kVinPx), (unsigned)nz);
    if (f64)
        hipLaunchKernelGGL(k_vol_in_t<double>, grid, dim3(256), 0, st, V_col, (int)sp.X, (int)sp.Y, (int)p->n,
                           (int)sp.Lx, sp.XP, sp.YP, (int)p->T, (int)p->S, which == 0 ? sp.vol_d : sp.htf_d,
                           (u16 *)nullptr, (u16 *)nullptr, p->nf, mark);
    else
        hipLaunchKernelGGL(k_vol_in_t<float>, grid, dim3(256), 0, st, V_col, (int)sp.X, (int)sp.Y, (int)p->n,
                           (int)sp.Lx, sp.XP, sp.YP, (int)p->T, (int)p->S, which == 0 ? sp.vol : sp.htf,
                           sh ? sp.vol_hi : nullptr, sh ? sp.vol_lo : nullptr, p->nf, mark);
    SID_TRY_HIP(hipGetLastError());
    return SID_OK;
}

int launch_img_in(sid_psf *p, const float *O_col, hipStream_t st, unsigned long long mark)
{
    SensorPlan &sp = p->plan;
    const bool f64 = p->precision == SID_PREC_FP64;
    const bool sh = bf16_shadows(p);
    const int64_t total = sp.X * sp.Y;
    if (f64)
        hipLaunchKernelGGL(k_vol_in<double>, dim3(grid_for(total)), dim3(256), 0, st, O_col, (int)sp.X, (int)sp.Y, 1,
                           (int)p->n, sp.XP, sp.YP, (int)p->T, (int)p->S, sp.img_d, (u16 *)nullptr, (u16 *)nullptr,
                           p->nf + 1, mark);
    else
        hipLaunchKernelGGL(k_vol_in<float>, dim3(grid_for(total)), dim3(256), 0, st, O_col, (int)sp.X, (int)sp.Y, 1,
                           (int)p->n, sp.XP, sp.YP, (int)p->T, (int)p->S, sp.img, sh ? sp.img_hi : nullptr,
                           sh ? sp.img_lo : nullptr, p->nf + 1, mark);
    SID_TRY_HIP(hipGetLastError());
    return SID_OK;
}

int launch_vol_out(sid_psf *p, float *V_col, int64_t nz, hipStream_t st)
{
    SensorPlan &sp = p->plan;
    const int64_t total = sp.X * sp.Y * nz;
    if (total == 0) return SID_OK;
    if (p->precision == SID_PREC_FP64)
        hipLaunchKernelGGL(k_vol_out<double>, dim3(grid_for(total)), dim3(256), 0, st, sp.vol_d, (int)sp.X,
                           (int)sp.Y, (int)nz, (int)p->n, sp.XP, sp.YP, (int)p->T, (int)p->S, V_col);
    else
        hipLaunchKernelGGL(k_vol_out<float>, dim3(grid_for(total)), dim3(256), 0, st, sp.vol, (int)sp.X, (int)sp.Y,
                           (int)nz, (int)p->n, sp.XP, sp.YP, (int)p->T, (int)p->S, V_col);
    SID_TRY_HIP(hipGetLastError());
    return SID_OK;
}

// ---------------------------------------------------------------------------
// Self-test of the MFMA lane layouts assumed above.
// ---------------------------------------------------------------------------
typedef _Float16 st_h8 __attribute__((ext_vector_type(8)));

__global__ void k_selftest(const float *A, const float *B, float *C32, float *C16, float *C64, float *C16h)
{
    // A: 16 x 32 row-major, B: 32 x 16 row-major
    const int lane = threadIdx.x, l16 = lane & 15, g = lane >> 4;
    f4 c = {0, 0, 0, 0};
    for (int s = 0; s < 8; ++s)
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(A[l16 * 32 + 8 * g + s], B[(8 * g + s) * 16 + l16], c, 0, 0, 0);
    for (int i = 0; i < 4; ++i) C32[(4 * g + i) * 16 + l16] = c[i];
    bf8 a, b;
    for (int j = 0; j < 8; ++j) {
        a[j] = (__bf16)A[l16 * 32 + 8 * g + j];
        b[j] = (__bf16)B[(8 * g + j) * 16 + l16];
    }
    f4 d = {0, 0, 0, 0};
    d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, d, 0, 0, 0);
    for (int i = 0; i < 4; ++i) C16[(4 * g + i) * 16 + l16] = d[i];
    st_h8 ah, bh;   // v_mfma_f32_16x16x32_f16 (the tiled FFT path's GEMM): the same layout
    for (int j = 0; j < 8; ++j) {
        ah[j] = (_Float16)A[l16 * 32 + 8 * g + j];
        bh[j] = (_Float16)B[(8 * g + j) * 16 + l16];
    }
    f4 h = {0, 0, 0, 0};
    h = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, h, 0, 0, 0);
    for (int i = 0; i < 4; ++i) C16h[(4 * g + i) * 16 + l16] = h[i];
    d4 e = {0, 0, 0, 0};
    for (int s = 0; s < 8; ++s)
        e = __builtin_amdgcn_mfma_f64_16x16x4f64((double)A[l16 * 32 + 8 * g + s], (double)B[(8 * g + s) * 16 + l16], e,
                                                0, 0, 0);
    for (int i = 0; i < 4; ++i) C64[(4 * i + g) * 16 + l16] = (float)e[i];
}

int run_selftest()
{
    std::vector<float> A(16 * 32), B(32 * 16), C32(256), C16(256), C64(256), C16h(256);
    // small integers: exact in bf16 and in fp32 accumulation
    for (int i = 0; i < 16 * 32; ++i) A[i] = (float)((i * 7 + 3) % 11) - 5.0f;
    for (int i = 0; i < 32 * 16; ++i) B[i] = (float)((i * 5 + 1) % 13) - 6.0f;
    float *dA, *dB, *dC, *dD, *dE, *dF;
    SID_TRY_HIP(hipMalloc(&dA, A.size() * 4));
    SID_TRY_HIP(hipMalloc(&dB, B.size() * 4));
    SID_TRY_HIP(hipMalloc(&dC, 256 * 4));
    SID_TRY_HIP(hipMalloc(&dD, 256 * 4));
    SID_TRY_HIP(hipMalloc(&dE, 256 * 4));
    SID_TRY_HIP(hipMalloc(&dF, 256 * 4));
    SID_TRY_HIP(hipMemcpy(dA, A.data(), A.size() * 4, hipMemcpyHostToDevice));
    SID_TRY_HIP(hipMemcpy(dB, B.data(), B.size() * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_selftest, dim3(1), dim3(64), 0, 0, dA, dB, dC, dD, dE, dF);
    SID_TRY_HIP(hipGetLastError());
    SID_TRY_HIP(hipMemcpy(C32.data(), dC, 256 * 4, hipMemcpyDeviceToHost));
    SID_TRY_HIP(hipMemcpy(C16.data(), dD, 256 * 4, hipMemcpyDeviceToHost));
    SID_TRY_HIP(hipMemcpy(C64.data(), dE, 256 * 4, hipMemcpyDeviceToHost));
    SID_TRY_HIP(hipMemcpy(C16h.data(), dF, 256 * 4, hipMemcpyDeviceToHost));
    (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dC); (void)hipFree(dD); (void)hipFree(dE); (void)hipFree(dF);
    for (int i = 0; i < 16; ++i)
        for (int j = 0; j < 16; ++j) {
            double s = 0;
            for (int k = 0; k < 32; ++k) s += (double)A[i * 32 + k] * B[k * 16 + j];
            if (C32[i * 16 + j] != (float)s || C16[i * 16 + j] != (float)s || C64[i * 16 + j] != (float)s ||
                C16h[i * 16 + j] != (float)s) {
                set_error("MFMA layout self-test failed at (%d,%d): fp32 %g bf16 %g fp64 %g fp16 %g expected %g", i,
                          j, C32[i * 16 + j], C16[i * 16 + j], C64[i * 16 + j], C16h[i * 16 + j], s);
                return SID_ERR_HIP;
            }
        }
    return SID_OK;
}

}  // namespace sid
