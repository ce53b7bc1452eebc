// Tiled FFT path of the light-field projector (BF16X2 precision mode).
//
// In lenslet coordinates the projection of plane c is a multichannel 2-D
// convolution (external/LFrecon_util/forwardProjectGPU.m:16-28 in the form of
// sid_direct.hip):
//     O[P, uv] = sum_{c, ab} sum_D V_c[P - D, ab] W_c[D][ab][uv],
//     W_c[D][ab][uv] = H_{a,b,c}[n D + (u - a, v - b) + off],
// D over the plane's lenslet taps.  With the taps of every plane on this path
// inside one range [Dlo, Dhi] (K = Dhi - Dlo + 1 taps per axis) the output is
// cut into tiles of B x B lenslets, and each tile is one circular convolution
// of length T = B + K - 1 (overlap-save): the input window of the tile at
// origin P0 starts at P0 - Dhi, and output P0 + p is element p + K - 1 of
//     IDFT_T( sum_{c,ab} G_c[f][ab][uv] * DFT_T(window_{c,ab})[f] ),
//     G_c[f][ab][uv] = sum_{e in [0,K)^2} W_c[Dlo + e][ab][uv] e^{-2 pi i f.e / T}.
// For one frequency bin f that sum is a complex GEMM, [uv] x [(c, ab)] times
// [(c, ab)] x [tile].  The backward (backwardProjectGPU_new.m:16-26; for the
// planes of this path it is the adjoint) correlates image windows starting at
// Q0 + Dlo with the same kernels:
//     B_c[Q0 + p, ab] = IDFT_T( sum_uv conj(G_c[f][ab][uv]) * DFT_T(image window_uv)[f] )[p].
// The data are real, so only fy in [0, T/2] is stored.
//
// The GEMMs run on MFMA (v_mfma_f32_16x16x32_f16) with FP32 accumulation:
// the spectra G in FP16, row (c, ab) scaled by 1/s_q (s_q the L1 norm of the
// kernel, so |G| <= 1); the input spectra split into FP16 hi + lo (together
// 21 bits), each tile's column scaled by a bound on its magnitudes (from the
// L1 norms of its windows) so that neither half can overflow.  G is stored once
// per direction in MFMA operand order: one wave's operand of a K step is one
// contiguous 1 KB run.  Against one GEMV over spectra of the whole padded
// sensor, the tiles shrink the spectra to (T / (L + K))^2 of their size and
// apply every stored value to all tiles at once.
//
// The window transforms are register FFTs (one sequence per thread, length T
// fixed at compile time, Stockham order with radices 4, 2, 3, 5) with one LDS
// transpose between the two axes.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "sid_fft_core.h"
#include "sid_internal.h"

namespace sid {

typedef unsigned int tu32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 th8 __attribute__((ext_vector_type(8)));

constexpr int kTfCG = 8;   // channels per transform block (4 real pairs)

// xcd_item / xcd_grid (sid_fft_core.h): the channel groups of one tile read
// and write the same 128-byte lines, so neighbouring items share an XCD.

// Register FFTs (reg_fft, sid_fft_core.h) transform the windows.

// Threads of a transform block: both phases in one pass where they fit.
template <int T>
constexpr int tf_threads()
{
    const int need = std::max(kTfCG / 2 * T, kTfCG * (T / 2 + 1));
    return std::min(256, (need + 63) / 64 * 64);
}

// Geometry of one transform launch.  Lenslet (x, y) of plane c of the source
// (k_tf_in) or destination (k_tf_out) is at ((c XP + x + pad) YP + y + pad) S
// + channel.  Blocks are (slot, tile, channel group).
struct TfGeom {
    int n2, chp, Lx, Ly, B, nty, ntile, ntp, nslot, ngrp, nz_in;
    int pad, S;
    int64_t XP, YP;
    int off;            // in: window origin = tile origin + off; out: first valid element of a tile
    int kslot, kdim;    // in: channel ch of slot s is column k = s kslot + ch of a row of kdim
};

// ---------------------------------------------------------------------------
// Window transforms: per block one (slot, tile) and kTfCG channels.  The
// channels are transformed in real pairs a + i b (Z = DFT(a + i b) gives
// A(k) = (Z(k) + conj Z(-k)) / 2, B(k) = (Z(k) - conj Z(-k)) / (2i)): along y
// in registers (one row per thread), the stored half spectrum through LDS,
// then along x (one column per thread).  Output: out[bin][k / 16][tile][k % 16], scaled by
// s_q (qscale, forward) and vmax[tile][slot ngrp + grp] = max over the block's
// channels of s_q T max_x sum_y |w[x][y]|, a bound on every |out| it wrote (at
// least the window's L1 norm; a maximum, so independent of thread timing).
// ---------------------------------------------------------------------------
// Channel stride (float2) of the transform kernels' LDS spectra, n = rows x H
// padded to 2 mod 8.  A 64-bit LDS access serves 16 lanes per cycle over 16
// eight-byte slots; the 16 lanes of a row pass are 4 channel pairs x 4 rows
// (offsets 2 p CS + r H, all distinct mod 16 for odd H when 2 CS = 4 mod 16)
// and those of a column pass 8 channels x 2 bins (offsets ch CS + k, distinct
// when CS = 2 mod 8).
__host__ __device__ constexpr int tf_lds_cs(int n) { return n + ((2 - n % 8) + 8) % 8; }

template <int T>
__global__ __launch_bounds__(tf_threads<T>()) void k_tf_in(const float *__restrict__ src,
                                                          const int *__restrict__ slot_plane,
                                                          const float *__restrict__ qscale, TfGeom g,
                                                          float2 *__restrict__ out, float *__restrict__ vmax)
{
    constexpr int H = T / 2 + 1, CG = kTfCG, NPAIR = CG / 2, NT = tf_threads<T>(), CS = tf_lds_cs(T * H);
    extern __shared__ float2 tsm[];   // [CG][T][H], channel stride CS
    __shared__ unsigned int rmax[CG];   // per channel: largest row sum of |w| (float bits, >= 0)
    const int item = xcd_item(g.nslot * g.ntile * g.ngrp);
    if (item < 0) return;
    const int grp = item % g.ngrp, rest = item / g.ngrp;
    const int tile = rest % g.ntile, slot = rest / g.ntile;
    const int tx = tile / g.nty, ty = tile - tx * g.nty;
    const int x0 = tx * g.B + g.off, y0 = ty * g.B + g.off;
    const int c = slot_plane ? slot_plane[slot] : 0;
    const bool live = c < g.nz_in;
    const int ch0 = grp * CG;
    // the window's columns inside the sensor: y in [ylo, yhi) (the same for the whole block)
    const int ylo = max(0, -y0), yhi = min(T, g.Ly - y0);
    if (threadIdx.x < CG) rmax[threadIdx.x] = 0u;
    __syncthreads();
    const float *base = src + (int64_t)c * g.XP * g.YP * g.S;
    // this thread's largest row sums of |w| (channels 2 pair, 2 pair + 1) as float bits: the
    // unsigned order of non-negative floats, with NaN above Inf (a NaN row poisons the bound)
    unsigned int ma = 0u, mb = 0u;
    for (int it = threadIdx.x; it < NPAIR * T; it += NT) {
        const int pair = it % NPAIR, x = it / NPAIR;
        const int ch = ch0 + 2 * pair;
        const int X = x0 + x;
        const bool rowin = live && X >= 0 && X < g.Lx && ch < g.n2;
        const bool two = ch + 1 < g.n2;
        // rows outside the sensor read a row inside it and are zeroed; the
        // second channel of a pair may be the record's padding (masked)
        const int Xc = min(max(X, 0), g.Lx - 1);
        const float *prow = base + ((int64_t)(Xc + g.pad) * g.YP + (y0 + g.pad)) * g.S + min(ch, g.n2 - 1);
        float2 z[T];
        float sa = 0.f, sb = 0.f;
#pragma unroll
        for (int y = 0; y < T; ++y) {
            float2 v = make_float2(0.f, 0.f);
            if (y >= ylo && y < yhi) v = ld_f2_a4(prow + (int64_t)y * g.S);
            v.x = rowin ? v.x : 0.f;
            v.y = rowin && two ? v.y : 0.f;
            z[y] = v;
            sa += fabsf(v.x);
            sb += fabsf(v.y);
        }
        reg_fft<T, -1>(z);
        float2 *dA = tsm + (2 * pair) * CS + x * H, *dB = tsm + (2 * pair + 1) * CS + x * H;
#pragma unroll
        for (int k = 0; k < H; ++k) {
            const float2 zk = z[k], zm = z[(T - k) % T];
            dA[k] = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
            dB[k] = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
        }
        ma = max(ma, __float_as_uint(sa));
        mb = max(mb, __float_as_uint(sb));
    }
    // lanes with equal lane % NPAIR hold the same pair (NT and 64 are multiples
    // of NPAIR): reduce over the wave, then one LDS atomic per pair and wave
    static_assert(NPAIR == 4, "the row-sum reduction assumes 4 pairs per block");
#pragma unroll
    for (int o = 4; o < 64; o <<= 1) {
        ma = max(ma, (unsigned int)__shfl_xor((int)ma, o));
        mb = max(mb, (unsigned int)__shfl_xor((int)mb, o));
    }
    if ((threadIdx.x & 63) < NPAIR) {
        const int pair = threadIdx.x & 3;
        atomicMax(&rmax[2 * pair], ma);
        atomicMax(&rmax[2 * pair + 1], mb);
    }
    __syncthreads();
    const int64_t bstride = (int64_t)g.ntile * g.kdim, fstep = (int64_t)H * bstride;
    for (int it = threadIdx.x; it < CG * H; it += NT) {
        const int ch = it % CG, k = it / CG;
        const int chg = ch0 + ch;
        float2 w[T];
        const float2 *s = tsm + ch * CS + k;
#pragma unroll
        for (int x = 0; x < T; ++x) w[x] = s[x * H];
        reg_fft<T, -1>(w);
        if (chg < g.n2) {
            const float q = qscale ? qscale[slot * g.n2 + chg] : 1.0f;
            const int kk = slot * g.kslot + chg;   // GEMM column, stored by K steps of 16 (all tiles of a step together)
            float2 *o = out + k * bstride + ((int64_t)(kk >> 4) * g.ntile + tile) * 16 + (kk & 15);
#pragma unroll
            for (int fx = 0; fx < T; ++fx) o[fx * fstep] = make_float2(w[fx].x * q, w[fx].y * q);
        }
    }
    if (threadIdx.x == 0) {
        float m = 0.f;
        for (int ch = 0; ch < CG; ++ch)
            if (ch0 + ch < g.n2)
                m = fmaxf(m, (float)T * __uint_as_float(rmax[ch]) * (qscale ? qscale[slot * g.n2 + ch0 + ch] : 1.0f));
        vmax[(int64_t)tile * g.nslot * g.ngrp + slot * g.ngrp + grp] = m;
    }
}

// The reduction that follows a backward projection, applied by the inverse
// transform itself when every plane of H is on the tiled path (no direct
// slabs to add): mode 1 writes the back-projection to the column-major output,
// mode 2 is the ISRA update of deconvRL.m:52,85-87 (r = Htf / d, Inf -> 0,
// X = X r, NaN -> 0) on the padded iterate, with the iterate optionally copied
// to the column-major output (the history).  A non-finite image stamp
// (check) makes d NaN everywhere; a non-finite new iterate sets the next stamp
// (mark).  The same arithmetic as k_bwd_reduce.
struct TfTail {
    int mode = 0;                       // 0: write the destination (no tail)
    int X = 0, Y = 0, n = 0;
    const float *htf = nullptr;         // padded layout, as the destination
    float *vol = nullptr;
    float *B_col = nullptr;             // column-major [X][Y][Nz], or null
    unsigned long long *nf = nullptr;
    unsigned long long check = 0, mark = 0;
};

// One element: plane c, lenslet (Px, Py), channel ab at padded offset o; true
// when the new iterate is not finite.
__device__ __forceinline__ bool tf_tail_elem(const TfTail &tl, int64_t o, int c, int Px, int Py, int ab, float d)
{
    const int x = Px * tl.n + ab % tl.n, y = Py * tl.n + ab / tl.n;
    const bool real = x < tl.X && y < tl.Y;
    const int64_t col = (int64_t)x + (int64_t)tl.X * ((int64_t)y + (int64_t)tl.Y * c);
    if (tl.mode == 1) {
        if (real) tl.B_col[col] = d;
        return false;
    }
    float v = 0.f;
    if (real) {
        float ratio = tl.htf[o] / d;
        if (isinf(ratio)) ratio = 0.f;
        v = tl.vol[o] * ratio;
        if (isnan(v)) v = 0.f;
    }
    tl.vol[o] = v;
    if (tl.B_col && real) tl.B_col[col] = v;
    return !isfinite(v);
}

// Inverse: spec[bin][slot][ntp][chp] (the sum of nsum partial spectra
// pstride apart) -> along x (one stored fy per thread; only the tile's valid
// rows kept in LDS), then along y in real pairs with the Hermitian completion;
// the valid B x B block, times scale * s_q (qscale, when given), goes to the
// destination layout (added to what is there when acc is set), or goes
// through the backward reduction's tail (tl.mode != 0).
template <int T, bool SUM>
__global__ __launch_bounds__(tf_threads<T>()) void k_tf_out(const float2 *__restrict__ spec, int nsum,
                                                           int64_t pstride, const int *__restrict__ slot_plane,
                                                           const float *__restrict__ qscale, TfGeom g, float scale,
                                                           int acc, float *__restrict__ dst, TfTail tl)
{
    constexpr int H = T / 2 + 1, CG = kTfCG, NPAIR = CG / 2, NT = tf_threads<T>();
    extern __shared__ float2 tsm[];   // [CG][B][H], channel stride CS
    const int B = g.B, CS = tf_lds_cs(B * H);
    const int item = xcd_item(g.nslot * g.ntile * g.ngrp);
    if (item < 0) return;
    const int grp = item % g.ngrp, rest = item / g.ngrp;
    const int tile = rest % g.ntile, slot = rest / g.ntile;
    const int tx = tile / g.nty, ty = tile - tx * g.nty;
    const int c = slot_plane ? slot_plane[slot] : 0;
    const int ch0 = grp * CG;
    const int64_t bstride = (int64_t)g.nslot * g.ntp * g.chp, fstep = (int64_t)H * bstride;
    for (int it = threadIdx.x; it < CG * H; it += NT) {
        const int ch = it % CG, k = it / CG;
        const int chg = ch0 + ch;
        const bool cin = chg < g.n2;   // channels up to the group's end are inside the chp padding
        float2 w[T];
        const float2 *s = spec + k * bstride + ((int64_t)slot * g.ntp + tile) * g.chp + chg;
#pragma unroll
        for (int fx = 0; fx < T; ++fx) {
            const float2 v = s[fx * fstep];
            w[fx] = cin ? v : make_float2(0.f, 0.f);
        }
        if constexpr (SUM) {
            for (int q = 1; q < nsum; ++q) {   // partial sums of a split GEMM, in a fixed order
                const float2 *sq = s + q * pstride;
#pragma unroll
                for (int fx = 0; fx < T; ++fx) {
                    const float2 v = sq[fx * fstep];
                    w[fx].x += cin ? v.x : 0.f;
                    w[fx].y += cin ? v.y : 0.f;
                }
            }
        }
        reg_fft<T, 1>(w);
#pragma unroll
        for (int x = 0; x < T; ++x) {
            const int xr = x - g.off;
            if (xr >= 0 && xr < B) tsm[ch * CS + xr * H + k] = w[x];
        }
    }
    __syncthreads();
    for (int it = threadIdx.x; it < NPAIR * B; it += NT) {
        const int pair = it % NPAIR, xr = it / NPAIR;
        const int Px = tx * B + xr;
        const int ch = ch0 + 2 * pair;
        if (Px >= g.Lx || ch >= g.n2) continue;
        const float2 *sA = tsm + (2 * pair) * CS + xr * H, *sB = tsm + (2 * pair + 1) * CS + xr * H;
        float2 z[T];
#pragma unroll
        for (int fy = 0; fy < T; ++fy) {
            float2 a, b;
            if (fy < H) {
                a = sA[fy];
                b = sB[fy];
            } else {
                const float2 ta = sA[T - fy], tb = sB[T - fy];
                a = make_float2(ta.x, -ta.y);
                b = make_float2(tb.x, -tb.y);
            }
            z[fy] = make_float2(a.x - b.y, a.y + b.x);
        }
        reg_fft<T, 1>(z);
        const bool two = ch + 1 < g.n2;
        const float qa = scale * (qscale ? qscale[slot * g.n2 + ch] : 1.0f);
        const float qb = two ? scale * (qscale ? qscale[slot * g.n2 + ch + 1] : 1.0f) : 0.f;
        const int64_t o = (((int64_t)c * g.XP + Px + g.pad) * g.YP + g.pad) * g.S + ch;
        if (tl.mode != 0) {
            const bool poison = tl.check != 0 && tl.nf[1] == tl.check;   // non-finite image: NaN everywhere
            bool bad = false;
#pragma unroll
            for (int y = 0; y < T; ++y) {
                const int yr = y - g.off, Py = ty * B + yr;
                if (yr >= 0 && yr < B && Py < g.Ly) {
                    const int64_t q = o + (int64_t)Py * g.S;
                    const float da = poison ? NAN : z[y].x * qa;
                    bad |= tf_tail_elem(tl, q, c, Px, Py, ch, da);
                    if (two) bad |= tf_tail_elem(tl, q + 1, c, Px, Py, ch + 1, poison ? NAN : z[y].y * qb);
                }
            }
            if (bad && tl.mark) tl.nf[0] = tl.mark;   // an Inf iterate poisons the next forward
            continue;
        }
#pragma unroll
        for (int y = 0; y < T; ++y) {
            const int yr = y - g.off, Py = ty * B + yr;
            if (yr >= 0 && yr < B && Py < g.Ly) {
                float *q = dst + o + (int64_t)Py * g.S;
                if (two) {
                    float2 v = make_float2(z[y].x * qa, z[y].y * qb);
                    if (acc) {
                        const float2 u = ld_f2_a4(q);
                        v.x += u.x;
                        v.y += u.y;
                    }
                    st_f2_a4(q, v);
                } else {
                    q[0] = acc ? q[0] + z[y].x * qa : z[y].x * qa;
                }
            }
        }
    }
}

// Per tile: the bound over its transform blocks (one wave per tile).
__global__ __launch_bounds__(256) void k_tf_tmax(const float *__restrict__ vmax, int ntile, int rowlen,
                                                 float *__restrict__ tmax)
{
    const int lane = threadIdx.x & 63, wpb = blockDim.x >> 6;
    for (int t = blockIdx.x * wpb + (threadIdx.x >> 6); t < ntile; t += gridDim.x * wpb) {
        float m = 0.f;
        for (int j = lane; j < rowlen; j += 64) m = fmaxf(m, vmax[(int64_t)t * rowlen + j]);
        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
        if (lane == 0) tmax[t] = m;
    }
}

// ---------------------------------------------------------------------------
// Spectra GEMM.  Block = one frequency bin and one part of its work; over all
// tiles it computes
//     C[row][tile] = sum_k A[row][k] * Bm[tile][k]        (complex)
// forward: rows uv, K = (plane, ab), one slot, the K steps cut into nsplit
// runs of `chunk` steps, part q writing its partial sum at out + q pstride;
// backward: for each of the planes in turn, rows ab, K = uv (16 K steps of 16
// per slot, uv padded), the slots cut into nsplit runs of `chunk` slots.  (684
// bins on 512 resident blocks would leave a second round two-thirds empty.)
// A: FP16 (re, im) pairs in fragment order, the 16 bytes of lane (l16, g) of
// flattened K step s (slot * KS + ks) and row block rb at A[((bin nslot KS + s)
// RB + rb) 64 + lane]: row rb*16 + l16, complex k = ks*16 + 4g .. 4g + 3.
// Bm: FP32 complex [bin][kB / 16][tile][16] (a K step of every tile is one
// contiguous run), scaled per tile by 1/tmax and split into
// FP16 hi + lo; Re C = A . (Br, -Bi), Im C = A . (Bi, Br) over the interleaved
// real K (v_mfma_f32_16x16x32_f16: lane (l16, g) feeds A[row l16][k 8g..8g+7],
// B[k 8g..8g+7][col l16] and holds C[4g + i][l16]).
// Pipeline: every wave keeps kTfRing K steps of its A fragments in flight (a
// register ring, refilled right after use); each thread converts its 32 bytes
// of Bm per step into the four operands in an LDS double buffer (one barrier
// per step), loaded kTfRing steps ahead.  The loop body has no branch (row
// blocks past RB repeat the last one, tiles past ntile read zeros and write to
// the padding columns of `out`), so the compiler's waits count the ring
// exactly.  Output out[bin][slot][ntp][chp] (ntp = tiles padded to 16) times
// tmax[tile].
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t pk_rtz(float a, float b)
{
    return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(a, b));
}
__device__ __forceinline__ float h_lo(uint32_t v) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(v & 0xffffu)); }
__device__ __forceinline__ float h_hi(uint32_t v) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(v >> 16)); }

// Four complex values (v0 = k, k+1; v1 = k+2, k+3) times s: the operands
// (re, -im) and (im, re), each as FP16 hi and lo (round toward zero: the
// halves of (im, re) are those of (re, -im), swapped and negated, exactly).
__device__ __forceinline__ void tf_split(float4 v0, float4 v1, float s, tu32x4 &rh, tu32x4 &rl, tu32x4 &ih,
                                         tu32x4 &il)
{
    const float re[4] = {v0.x * s, v0.z * s, v1.x * s, v1.z * s};
    const float im[4] = {v0.y * s, v0.w * s, v1.y * s, v1.w * s};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t h = pk_rtz(re[j], -im[j]);
        const float r0 = re[j] - h_lo(h), r1 = -im[j] - h_hi(h);
        rh[j] = h;
        rl[j] = pk_rtz(r0, r1);
        ih[j] = pk_rtz(im[j], re[j]);
        il[j] = pk_rtz(-r1, r0);
    }
}

__device__ __forceinline__ f4 mfma16(tu32x4 a, tu32x4 b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(th8, a), __builtin_bit_cast(th8, b), c, 0, 0, 0);
}

constexpr int kTfRing = 4;    // K steps in flight per wave
constexpr int kTfKsB = 16;    // backward K steps per slot (uv padded to 256)

template <int NTB, bool BWD>
__global__ __launch_bounds__(256, 2) void k_tf_gemm(const tu32x4 *__restrict__ A, const float2 *__restrict__ Bm,
                                                    const float *__restrict__ tmax, int nbin, int nslot, int ntile,
                                                    int ntp, int KS, int kB, int RB, int chp, int nsplit, int chunk,
                                                    int64_t pstride, float2 *__restrict__ out)
{
    constexpr int D = NTB >= 4 ? 2 : kTfRing;   // K steps in flight (four column blocks leave room for two)
    constexpr int U = D;   // steps per loop trip
    static_assert(U % 2 == 0 && kTfKsB % U == 0, "LDS buffers repeat within a trip, slots end at trip ends");
    __shared__ tu32x4 bop[2][4][4][64];   // [buffer][operand rh, rl, ih, il][tile block][slot]
    const int item = xcd_item(nbin * nsplit);
    if (item < 0) return;
    const int bin = item / nsplit, part = item - bin * nsplit;
    // this block's steps: forward a run of K steps, backward a run of slots
    int s_lo, nsteps, slot0 = 0;
    if constexpr (BWD) {
        slot0 = part * chunk;
        s_lo = slot0 * KS;
        nsteps = min(chunk, nslot - slot0) * KS;
    } else {
        s_lo = part * chunk;
        nsteps = min(chunk, nslot * KS - s_lo);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, l16 = lane & 15, g = lane >> 4;
    // this thread's Bm chunk of every step: tile tt, complex 4 qt .. 4 qt + 3
    const int tt = threadIdx.x >> 2, qt = threadIdx.x & 3;
    const bool tvalid = tt < ntile;
    const float mt = tvalid ? tmax[tt] : 0.f;
    const float sc = mt > 0.f ? 1.0f / mt : 0.f;
    const int64_t bstep = (int64_t)ntile * 16;   // one K step of every tile
    const float2 *bsrc = Bm + (int64_t)bin * ntile * kB + (tvalid ? tt : 0) * 16 + 4 * qt + (BWD ? 0 : s_lo * bstep);
    // 16-byte slot of (column c, k group g) in a tile block: 16 g + ((c + 4 g) & 15),
    // so the 16 lanes of each quarter wave write and read 16 different bank groups
    tu32x4 *bdst = &bop[0][0][tt >> 4][16 * qt + (((tt & 15) + 4 * qt) & 15)];
    const float4 zf = make_float4(0.f, 0.f, 0.f, 0.f);
    float mx[NTB];
#pragma unroll
    for (int tb = 0; tb < NTB; ++tb) mx[tb] = tb * 16 + l16 < ntile ? tmax[tb * 16 + l16] : 0.f;
    int rbw[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) rbw[i] = min(wave + 4 * i, RB - 1);
    const int64_t astep = (int64_t)RB * 64;
    const tu32x4 *abase = A + ((int64_t)bin * nslot * KS + s_lo) * astep + lane;
    tu32x4 a[D][4];
    float4 b0[D], b1[D];
    // the ring's first D steps, loaded in the loop's own order (B, then A):
    // the loads pending at the loop's entry then match those pending on its
    // back edge, so the compiler's waits at the top of a trip count the ring
    // instead of draining every load in flight
#pragma unroll
    for (int r = 0; r < D; ++r) {
        const int s = min(r, nsteps - 1);
        const float2 *pb = bsrc + (BWD ? (s & (kTfKsB - 1)) : s) * bstep;
        b0[r] = *(const float4 *)pb;
        b1[r] = *(const float4 *)(pb + 2);
        const tu32x4 *pa = abase + (int64_t)s * astep;
#pragma unroll
        for (int i = 0; i < 4; ++i) a[r][i] = __builtin_nontemporal_load(pa + rbw[i] * 64);
        __builtin_amdgcn_sched_barrier(0);   // keep that order
    }
    f4 cr[4][NTB], ci[4][NTB];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int tb = 0; tb < NTB; ++tb) cr[i][tb] = ci[i][tb] = f4{0.f, 0.f, 0.f, 0.f};
    for (int s0 = 0; s0 < nsteps; s0 += U) {
#pragma unroll
        for (int j = 0; j < U; ++j) {
            const int r = j % D;
            const int sn = min(s0 + j + D, nsteps - 1);
            {
                tu32x4 rh, rl, ih, il;
                tf_split(tvalid ? b0[r] : zf, tvalid ? b1[r] : zf, sc, rh, rl, ih, il);
                tu32x4 *d = bdst + (j & 1) * (4 * 4 * 64);
                d[0] = rh;
                d[4 * 64] = rl;
                d[2 * 4 * 64] = ih;
                d[3 * 4 * 64] = il;
            }
            {
                const float2 *pb = bsrc + (BWD ? (sn & (kTfKsB - 1)) : sn) * bstep;
                b0[r] = *(const float4 *)pb;
                b1[r] = *(const float4 *)(pb + 2);
            }
            __syncthreads();
            const tu32x4 *bs = &bop[j & 1][0][0][16 * g + ((l16 + 4 * g) & 15)];
#pragma unroll
            for (int tb = 0; tb < NTB; ++tb) {
                const tu32x4 rh = bs[tb * 64], rl = bs[(4 + tb) * 64], ih = bs[(8 + tb) * 64], il = bs[(12 + tb) * 64];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    cr[i][tb] = mfma16(a[r][i], rh, cr[i][tb]);
                    ci[i][tb] = mfma16(a[r][i], ih, ci[i][tb]);
                    cr[i][tb] = mfma16(a[r][i], rl, cr[i][tb]);
                    ci[i][tb] = mfma16(a[r][i], il, ci[i][tb]);
                }
            }
            const tu32x4 *pa = abase + (int64_t)sn * astep;
#pragma unroll
            for (int i = 0; i < 4; ++i) a[r][i] = __builtin_nontemporal_load(pa + rbw[i] * 64);
            __builtin_amdgcn_sched_barrier(0);   // keep the steps apart (register pressure)
        }
        if (BWD && ((s0 + U) & (kTfKsB - 1)) == 0) {
            // slot slot0 + s0 / 16 is complete: write it
            float2 *ob = out + ((int64_t)bin * nslot + slot0 + s0 / kTfKsB) * ntp * chp + 4 * g;
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int tb = 0; tb < NTB; ++tb) {
                    // a repeated row block (RB = 15), or a padding tile column: nothing to store
                    if (wave + 4 * i >= RB || tb * 16 + l16 >= ntile) continue;
                    float2 *o = ob + (int64_t)(tb * 16 + l16) * chp + rbw[i] * 16;
                    const float m = mx[tb];
                    *(float4 *)o = make_float4(cr[i][tb][0] * m, ci[i][tb][0] * m, cr[i][tb][1] * m, ci[i][tb][1] * m);
                    *(float4 *)(o + 2) =
                        make_float4(cr[i][tb][2] * m, ci[i][tb][2] * m, cr[i][tb][3] * m, ci[i][tb][3] * m);
                    cr[i][tb] = ci[i][tb] = f4{0.f, 0.f, 0.f, 0.f};
                }
        }
    }
    if constexpr (!BWD) {
        float2 *ob = out + part * pstride + (int64_t)bin * ntp * chp + 4 * g;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int tb = 0; tb < NTB; ++tb) {
                if (wave + 4 * i >= RB || tb * 16 + l16 >= ntile) continue;   // as in the backward
                float2 *o = ob + (int64_t)(tb * 16 + l16) * chp + rbw[i] * 16;
                const float m = mx[tb];
                *(float4 *)o = make_float4(cr[i][tb][0] * m, ci[i][tb][0] * m, cr[i][tb][1] * m, ci[i][tb][1] * m);
                *(float4 *)(o + 2) = make_float4(cr[i][tb][2] * m, ci[i][tb][2] * m, cr[i][tb][3] * m, ci[i][tb][3] * m);
            }
    }
}

// ---------------------------------------------------------------------------
// Spectra build, one slot (plane) at a time.
//   k_tf_g1: X1[ab][ex][fy][uv] = sum_ey W[Dlo + ex, Dlo + ey][ab][uv] e^{-2 pi i fy ey / T}
//   k_tf_g2: Gt[fx][fy][ab][uv] = sum_ex X1[ab][ex][fy][uv] e^{-2 pi i fx ex / T}
//   k_tf_sf / k_tf_sb: Gt / s_q in FP16 into the forward / backward operand order
// ---------------------------------------------------------------------------
template <int T>
__global__ __launch_bounds__(256) void k_tf_g1(const float *__restrict__ Hm, int64_t M, int n, int c, int lo, int hi,
                                               int off, int Dlo, int K, int chp, const float2 *__restrict__ tw,
                                               float2 *__restrict__ X1)
{
    constexpr int H = T / 2 + 1;
    const int ex = blockIdx.x % K, ab = blockIdx.x / K;
    const int uv = threadIdx.x;
    const int n2 = n * n;
    if (uv >= chp) return;
    float2 acc[H];
#pragma unroll
    for (int f = 0; f < H; ++f) acc[f] = make_float2(0.f, 0.f);
    const int a = ab % n, b = ab / n, u = uv % n, v = uv / n;
    const int i = n * (Dlo + ex) + u - a + off;
    if (uv < n2 && i >= lo && i <= hi) {
        const float *h = Hm + (int64_t)i + M * M * ((int64_t)a + n * ((int64_t)b + (int64_t)n * c));
#pragma unroll
        for (int ey = 0; ey < T; ++ey) {
            const int j = n * (Dlo + ey) + v - b + off;
            const float w = (ey < K && j >= lo && j <= hi) ? h[M * (int64_t)j] : 0.f;
#pragma unroll
            for (int f = 0; f < H; ++f) {
                const float2 t = tw[(f * ey) % T];
                acc[f].x += w * t.x;
                acc[f].y += w * t.y;
            }
        }
    }
    float2 *o = X1 + (((int64_t)ab * K + ex) * H) * chp + uv;
#pragma unroll
    for (int f = 0; f < H; ++f) o[(int64_t)f * chp] = acc[f];
}

template <int T>
__global__ __launch_bounds__(256) void k_tf_g2(const float2 *__restrict__ X1, int n2, int K, int chp,
                                               const float2 *__restrict__ tw, float2 *__restrict__ Gt)
{
    constexpr int H = T / 2 + 1;
    const int fy = blockIdx.x % H, ab = blockIdx.x / H;
    const int uv = threadIdx.x;
    if (uv >= chp) return;
    float2 x[T];
#pragma unroll
    for (int ex = 0; ex < T; ++ex)
        x[ex] = ex < K ? X1[(((int64_t)ab * K + ex) * H + fy) * chp + uv] : make_float2(0.f, 0.f);
#pragma unroll 1
    for (int fx = 0; fx < T; ++fx) {
        float2 s = make_float2(0.f, 0.f);
#pragma unroll
        for (int ex = 0; ex < T; ++ex) {
            const float2 t = tw[(fx * ex) % T];
            s.x += x[ex].x * t.x - x[ex].y * t.y;
            s.y += x[ex].x * t.y + x[ex].y * t.x;
        }
        Gt[(((int64_t)fx * H + fy) * n2 + ab) * chp + uv] = s;
    }
}

__device__ __forceinline__ uint32_t pk_rn(float a, float b)
{
    return __builtin_bit_cast(uint32_t, __floats2half2_rn(a, b));
}

// Forward operand of slot `slot`: row uv, K = (slot, ab) at k = slot kslot + ab.
__global__ void k_tf_sf(const float2 *__restrict__ Gt, const float *__restrict__ qscale, int nbin, int n2, int chp,
                        int kslot, int slot, int KS, int RB, tu32x4 *__restrict__ Gf)
{
    const int nq = kslot / 4;
    const int64_t total = (int64_t)nbin * nq * chp;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int uv = (int)(e % chp);
        const int64_t t = e / chp;
        const int kq = (int)(t % nq), bin = (int)(t / nq);
        const int kg = slot * nq + kq;   // global quad: k = 4 kg .. 4 kg + 3
        const int ks = kg >> 2, gq = kg & 3;
        tu32x4 u;
#pragma unroll
        for (int jp = 0; jp < 4; ++jp) {
            const int ab = 4 * kq + jp;
            uint32_t w = 0u;
            if (ab < n2 && uv < n2) {
                const float2 v = Gt[((int64_t)bin * n2 + ab) * chp + uv];
                const float s = 1.0f / qscale[slot * n2 + ab];
                w = pk_rn(v.x * s, v.y * s);
            }
            u[jp] = w;
        }
        Gf[(((int64_t)bin * KS + ks) * RB + (uv >> 4)) * 64 + (uv & 15) + 16 * gq] = u;
    }
}

// Backward operand of slot `slot`: row ab, K = uv (kTfKsB steps, zero past
// n2), conj(G) / s_q.
__global__ void k_tf_sb(const float2 *__restrict__ Gt, const float *__restrict__ qscale, int nbin, int n2, int chp,
                        int nslot, int slot, int RB, tu32x4 *__restrict__ Gb)
{
    const int nq = kTfKsB * 4;
    const int64_t total = (int64_t)nbin * nq * chp;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int ab = (int)(e % chp);
        const int64_t t = e / chp;
        const int uq = (int)(t % nq), bin = (int)(t / nq);
        const int ks = uq >> 2, gq = uq & 3;
        tu32x4 u;
        const float s = ab < n2 ? 1.0f / qscale[slot * n2 + ab] : 0.f;
#pragma unroll
        for (int jp = 0; jp < 4; ++jp) {
            const int uv = 4 * uq + jp;
            uint32_t w = 0u;
            if (ab < n2 && uv < n2) {
                const float2 v = Gt[((int64_t)bin * n2 + ab) * chp + uv];
                w = pk_rn(v.x * s, -v.y * s);
            }
            u[jp] = w;
        }
        Gb[((((int64_t)bin * nslot + slot) * kTfKsB + ks) * RB + (ab >> 4)) * 64 + (ab & 15) + 16 * gq] = u;
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
// FFT lengths with instantiated transforms.
static const int kTfLens[] = {16, 24, 32, 36, 40, 48, 64};

template <int T>
static int tf_set_lds()
{
    constexpr size_t lds = sizeof(float2) * kTfCG * tf_lds_cs(T * (T / 2 + 1));   // bounds both kernels' areas
    if (lds > 65536) {
        SID_TRY_HIP(hipFuncSetAttribute((const void *)k_tf_in<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        SID_TRY_HIP(hipFuncSetAttribute((const void *)k_tf_out<T, false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        SID_TRY_HIP(hipFuncSetAttribute((const void *)k_tf_out<T, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    return SID_OK;
}

template <int T>
static int tf_spectra(sid_psf *p, const TfGroup &g, hipStream_t st, float2 *X1, float2 *Gt)
{
    FftState &f = p->plan.fft;
    const int n = (int)p->n, n2 = n * n;
    const int RB = f.chp / 16, KSf = g.kpad / 16;
    const float *qs = f.qscale + (int64_t)g.s0 * n2;
    for (int sl = 0; sl < g.ns; ++sl) {
        const int c = f.planes[g.s0 + sl];
        const Plane &pl = p->planes[c];
        hipLaunchKernelGGL(k_tf_g1<T>, dim3(n2 * g.K), dim3(256), 0, st, p->H, p->M, n, c, pl.lo, pl.hi, pl.off_f,
                           g.Dlo, g.K, f.chp, g.ft.tw_fwd, X1);
        SID_TRY_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_tf_g2<T>, dim3(n2 * g.H), dim3(256), 0, st, X1, n2, g.K, f.chp, g.ft.tw_fwd, Gt);
        SID_TRY_HIP(hipGetLastError());
        const int64_t tf_ = (int64_t)g.nbin * (f.kslot / 4) * f.chp, tb_ = (int64_t)g.nbin * (kTfKsB * 4) * f.chp;
        hipLaunchKernelGGL(k_tf_sf, dim3((unsigned)std::min<int64_t>(65536, (tf_ + 255) / 256)), dim3(256), 0, st, Gt,
                           qs, g.nbin, n2, f.chp, f.kslot, sl, KSf, RB, (tu32x4 *)g.G);
        SID_TRY_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_tf_sb, dim3((unsigned)std::min<int64_t>(65536, (tb_ + 255) / 256)), dim3(256), 0, st, Gt,
                           qs, g.nbin, n2, f.chp, g.ns, sl, RB, (tu32x4 *)g.Gb);
        SID_TRY_HIP(hipGetLastError());
    }
    return SID_OK;
}

template <int T>
static int tf_dispatch_build(sid_psf *p, const TfGroup &g, hipStream_t st, float2 *X1, float2 *Gt)
{
    SID_TRY(tf_set_lds<T>());
    return tf_spectra<T>(p, g, st, X1, Gt);
}

static int tf_build_for(sid_psf *p, const TfGroup &g, hipStream_t st, float2 *X1, float2 *Gt)
{
    switch (g.T) {
    case 16: return tf_dispatch_build<16>(p, g, st, X1, Gt);
    case 24: return tf_dispatch_build<24>(p, g, st, X1, Gt);
    case 32: return tf_dispatch_build<32>(p, g, st, X1, Gt);
    case 36: return tf_dispatch_build<36>(p, g, st, X1, Gt);
    case 40: return tf_dispatch_build<40>(p, g, st, X1, Gt);
    case 48: return tf_dispatch_build<48>(p, g, st, X1, Gt);
    case 64: return tf_dispatch_build<64>(p, g, st, X1, Gt);
    default: set_error("internal: no tiled transform of length %d", g.T); return SID_ERR_ARG;
    }
}

// Modelled time of one group of P planes with common tap width K at
// transform length T, forward plus backward, in seconds: each GEMM the larger
// of its HBM bytes (spectra once; the input spectra or the backward result
// written and read) at 4.8 TB/s and its MFMA work at 1.2 PFLOP/s, the window
// transforms in proportion to planes x tiles x T^2 log2 T (6.8e-10 s per unit
// at T = 36, measured at 1005^2 x 51), and 40 us of launches.  The rates are
// the ones the kernels reach on MI355X; only their ratios steer the choice.
struct TfChoice {
    int T = 0, B = 0, ntx = 0, nty = 0;
    double cost = 0;
};

static TfChoice tf_choose(int K, int Lx, int Ly, int P, int chp, int kslot)
{
    TfChoice best;
    const int force = std::getenv("SID_TF_T") ? std::atoi(std::getenv("SID_TF_T")) : 0;
    // SID_TF_GROUP_US replaces the 40 us of launches per group (tests: groups on small problems)
    const double launch = std::getenv("SID_TF_GROUP_US") ? 1e-6 * std::atof(std::getenv("SID_TF_GROUP_US")) : 40e-6;
    for (int T : kTfLens) {
        if (force && T != force) continue;
        const int B = T - K + 1;
        if (B < 1) continue;
        const int ntx = (Lx + B - 1) / B, nty = (Ly + B - 1) / B, ntile = ntx * nty;
        const int ntb = (ntile + 15) / 16;
        if (ntb > 4) continue;
        const double nbin = (double)T * (T / 2 + 1), rows = chp;
        const double kf = (double)P * kslot, kb = (double)P * kTfKsB * 16;
        const double bytes_f = nbin * (rows * kf * 4 + 2.0 * ntile * kf * 8);
        const double bytes_b = nbin * (rows * kb * 4 + 2.0 * ntb * 16 * rows * P * 8);
        const double flops_f = nbin * 16.0 * ntb * rows * kf * 16, flops_b = nbin * 16.0 * ntb * rows * kb * 16;
        const double xform = 6.8e-10 * P * ntile * (double)T * T * std::log2((double)T) / std::log2(36.0);
        const double cost = std::max(bytes_f / 4.8e12, flops_f / 1.2e15) + std::max(bytes_b / 4.8e12, flops_b / 1.2e15) +
                            xform + launch;
        if (best.T == 0 || cost < best.cost) best = TfChoice{T, B, ntx, nty, cost};
    }
    return best;
}

int ensure_aux(sid_psf *p)
{
    if (p->aux) return SID_OK;
    SID_TRY_HIP(hipStreamCreateWithFlags(&p->aux, hipStreamNonBlocking));
    SID_TRY_HIP(hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming));
    SID_TRY_HIP(hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming));
    return SID_OK;
}

void tf_free(FftState &f)
{
    for (TfGroup &g : f.tg) {
        fft_plan_free(g.ft);
        for (void *q : {g.G, g.Gb, (void *)g.vin, (void *)g.zin, (void *)g.vmax, (void *)g.tmax})
            if (q) (void)hipFree(q);
    }
    f.tg.clear();
    for (void *q : {(void *)f.yout, (void *)f.bout, (void *)f.tout, (void *)f.yout2, (void *)f.bout2, (void *)f.tout2})
        if (q) (void)hipFree(q);
    f.yout = f.bout = f.yout2 = f.bout2 = nullptr;
    f.tout = f.tout2 = nullptr;
    for (hipEvent_t e : {f.tev_fork, f.tev_join})
        if (e) (void)hipEventDestroy(e);
    f.tev_fork = f.tev_join = nullptr;
    f.nchain = 1;
    f.tiled = false;
}

// Work splits of a group's two GEMMs.  One block per bin fills the GPU when
// there are at least as many bins as resident blocks (two per CU); a smaller
// group cuts each bin's work into parts so that bins x parts fills it once:
// the forward into runs of K steps (partial sums, added in a fixed order by the
// inverse transform), the backward into runs of slots.  More parts than that
// gained nothing at 1005^2 x 51 (the GEMMs stream HBM; a last round of blocks
// two-thirds empty runs faster per block) and the forward's partial sums cost
// their traffic.  SID_TF_SPLIT_F / SID_TF_SPLIT_B force the number of parts.
static void tf_splits(TfGroup &g, int cus)
{
    const int64_t res = 2 * (int64_t)cus;
    auto pick = [&](int units, int gran, const char *env, int &nsplit, int &chunk) {
        int want = (int)((res + g.nbin - 1) / g.nbin);
        if (const char *e = std::getenv(env)) want = std::max(1, std::atoi(e));
        want = std::max(1, std::min(want, std::max(1, units / gran)));
        chunk = ((units + want - 1) / want + gran - 1) / gran * gran;
        nsplit = (units + chunk - 1) / chunk;
    };
    pick(g.kpad / 16, kTfRing, "SID_TF_SPLIT_F", g.tsplit_f, g.tchunk_f);
    pick(g.ns, 1, "SID_TF_SPLIT_B", g.tsplit_b, g.tchunk_b);
}

// Geometry, buffers and spectra of the tiled form for the planes in f.planes,
// which fft_state_build orders by ascending tap width (f.qscale, f.dfft and the
// plane tables are already built).  The slots are cut into at most
// SID_TF_GROUPS (default 6) contiguous groups by dynamic programming over the
// modelled group times: a narrow plane in a group of wide ones pays for the
// wide group's transform length, a group of its own pays launches.
int tf_build(sid_psf *p, hipStream_t st)
{
    SensorPlan &sp = p->plan;
    FftState &f = sp.fft;
    const int n = (int)p->n, n2 = n * n, P = (int)f.planes.size();
    if (n2 > 256 || P == 0) return kTfNotApplicable;
    const int chp = (n2 + 15) / 16 * 16, kslot = (n2 + 3) / 4 * 4;
    const int Lx = (int)sp.Lx, Ly = (int)sp.Ly;
    int maxg = 6;
    if (const char *e = std::getenv("SID_TF_GROUPS")) maxg = std::max(1, std::atoi(e));
    // groups of slots [i, j): common tap range and modelled cost
    auto range = [&](int i, int j, int &lo, int &hi) {
        lo = INT_MAX;
        hi = INT_MIN;
        for (int q = i; q < j; ++q) {
            lo = std::min(lo, p->planes[f.planes[q]].tf_lo);
            hi = std::max(hi, p->planes[f.planes[q]].tf_hi);
        }
    };
    std::vector<std::vector<double>> gc(P + 1, std::vector<double>(P + 1, -1.0));
    for (int i = 0; i < P; ++i)
        for (int j = i + 1; j <= P; ++j) {
            int lo, hi;
            range(i, j, lo, hi);
            const TfChoice tc = tf_choose(hi - lo + 1, Lx, Ly, j - i, chp, kslot);
            gc[i][j] = tc.T ? tc.cost : -1.0;
        }
    // best[g][i]: cheapest cut of slots [i, P) into at most g groups
    const double inf = 1e30;
    std::vector<std::vector<double>> best(maxg + 1, std::vector<double>(P + 1, inf));
    std::vector<std::vector<int>> cut(maxg + 1, std::vector<int>(P + 1, -1));
    for (int g = 0; g <= maxg; ++g) best[g][P] = 0.0;
    for (int g = 1; g <= maxg; ++g)
        for (int i = P - 1; i >= 0; --i)
            for (int j = i + 1; j <= P; ++j) {
                if (gc[i][j] < 0 || best[g - 1][j] >= inf) continue;
                const double c = gc[i][j] + best[g - 1][j];
                if (c < best[g][i]) {
                    best[g][i] = c;
                    cut[g][i] = j;
                }
            }
    if (best[maxg][0] >= inf) return kTfNotApplicable;
    int cus = 256;
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, p->device) == hipSuccess) cus = prop.multiProcessorCount;
    }
    f.chp = chp;
    f.kslot = kslot;
    f.ngrp = (n2 + kTfCG - 1) / kTfCG;
    f.tiled = true;
    int64_t nyout = 0, nbout = 0;
    for (int g = maxg, i = 0; i < P; i = cut[g][i], --g) {
        TfGroup t;
        t.s0 = i;
        t.ns = cut[g][i] - i;
        range(i, cut[g][i], t.Dlo, t.Dhi);
        t.K = t.Dhi - t.Dlo + 1;
        const TfChoice tc = tf_choose(t.K, Lx, Ly, t.ns, chp, kslot);
        t.T = tc.T;
        t.H = tc.T / 2 + 1;
        t.B = tc.B;
        t.ntx = tc.ntx;
        t.nty = tc.nty;
        t.ntile = tc.ntx * tc.nty;
        t.nbin = t.T * t.H;
        t.kpad = (t.ns * kslot + 16 * kTfRing - 1) / (16 * kTfRing) * (16 * kTfRing);   // whole ring trips
        tf_splits(t, cus);
        t.cost = tc.cost;
        f.tg.push_back(t);
    }
    // Two chains when every plane is on this path and there are two groups or
    // more: the groups, in order, go to the chain with the smaller modelled
    // time so far; the chains run side by side on the caller's stream and the
    // handle's second stream, filling each other's launch gaps and tails.
    f.nchain = (p->overlap && f.all_fft && f.tg.size() >= 2) ? 2 : 1;
    {
        double load[2] = {0.0, 0.0};
        for (TfGroup &t : f.tg) {
            t.chain = f.nchain == 2 && load[1] < load[0] ? 1 : 0;
            load[t.chain] += t.cost;
        }
    }
    int64_t nyout2 = 0, nbout2 = 0;
    for (const TfGroup &t : f.tg) {
        const int64_t ntp = (t.ntile + 15) / 16 * 16;
        const int64_t y = (int64_t)t.nbin * ntp * chp * t.tsplit_f, b = (int64_t)t.nbin * ntp * chp * t.ns;
        (t.chain ? nyout2 : nyout) = std::max(t.chain ? nyout2 : nyout, y);
        (t.chain ? nbout2 : nbout) = std::max(t.chain ? nbout2 : nbout, b);
    }
    const int64_t nimg = sp.Lx * sp.Ly * p->NP;
    SID_TRY_HIP(hipMalloc(&f.yout, sizeof(float2) * nyout));
    SID_TRY_HIP(hipMalloc(&f.bout, sizeof(float2) * nbout));
    SID_TRY_HIP(hipMalloc(&f.tout, sizeof(float) * nimg));
    SID_TRY_HIP(hipMemsetAsync(f.tout, 0, sizeof(float) * nimg, st));
    if (f.nchain == 2) {
        SID_TRY(ensure_aux(p));
        SID_TRY_HIP(hipMalloc(&f.yout2, sizeof(float2) * nyout2));
        SID_TRY_HIP(hipMalloc(&f.bout2, sizeof(float2) * nbout2));
        SID_TRY_HIP(hipMalloc(&f.tout2, sizeof(float) * nimg));
        SID_TRY_HIP(hipMemsetAsync(f.tout2, 0, sizeof(float) * nimg, st));
        SID_TRY_HIP(hipEventCreateWithFlags(&f.tev_fork, hipEventDisableTiming));
        SID_TRY_HIP(hipEventCreateWithFlags(&f.tev_join, hipEventDisableTiming));
    }
    f.g_bytes = 0;
    int64_t x1_elems = 0, gt_elems = 0;
    const int RB = chp / 16;
    for (TfGroup &t : f.tg) {
        SID_TRY(fft_plan_make(t.ft, t.T));
        const int64_t gf = (int64_t)t.nbin * (t.kpad / 16) * RB * 64 * 16,
                      gb = (int64_t)t.nbin * t.ns * kTfKsB * RB * 64 * 16;
        SID_TRY_HIP(hipMalloc(&t.G, gf));
        SID_TRY_HIP(hipMalloc(&t.Gb, gb));
        SID_TRY_HIP(hipMemsetAsync(t.G, 0, gf, st));
        SID_TRY_HIP(hipMemsetAsync(t.Gb, 0, gb, st));
        f.g_bytes += gf + gb;
        const int64_t nvin = (int64_t)t.nbin * t.ntile * t.kpad, nzin = (int64_t)t.nbin * t.ntile * kTfKsB * 16;
        SID_TRY_HIP(hipMalloc(&t.vin, sizeof(float2) * nvin));
        SID_TRY_HIP(hipMemsetAsync(t.vin, 0, sizeof(float2) * nvin, st));   // K padding stays zero
        SID_TRY_HIP(hipMalloc(&t.zin, sizeof(float2) * nzin));
        SID_TRY_HIP(hipMemsetAsync(t.zin, 0, sizeof(float2) * nzin, st));   // uv padding stays zero
        SID_TRY_HIP(hipMalloc(&t.vmax, sizeof(float) * (int64_t)t.ntile * t.ns * f.ngrp));
        SID_TRY_HIP(hipMalloc(&t.tmax, sizeof(float) * t.ntile));
        x1_elems = std::max<int64_t>(x1_elems, (int64_t)n2 * t.K * t.H * chp);
        gt_elems = std::max<int64_t>(gt_elems, (int64_t)t.nbin * n2 * chp);
    }
    // spectra, slot by slot through two scratch buffers
    float2 *X1 = nullptr, *Gt = nullptr;
    SID_TRY_HIP(hipMalloc(&X1, sizeof(float2) * x1_elems));
    if (hipMalloc(&Gt, sizeof(float2) * gt_elems) != hipSuccess) {
        (void)hipFree(X1);
        set_error("hipMalloc of the tiled spectra scratch failed");
        return SID_ERR_OOM;
    }
    int rc = SID_OK;
    for (const TfGroup &t : f.tg) {
        rc = tf_build_for(p, t, st, X1, Gt);
        if (rc != SID_OK) break;
    }
    const hipError_t e = hipStreamSynchronize(st);
    (void)hipFree(X1);
    (void)hipFree(Gt);
    if (rc != SID_OK) return rc;
    if (e != hipSuccess) {
        set_error("tiled spectra build failed: %s", hipGetErrorString(e));
        return SID_ERR_HIP;
    }
    return SID_OK;
}

void tf_gemm_work(const FftState &f, int64_t bytes[2], double flops[2])
{
    bytes[0] = bytes[1] = 0;
    flops[0] = flops[1] = 0.0;
    const int RB = f.chp / 16;
    for (const TfGroup &t : f.tg) {
        const int64_t ntb = (t.ntile + 15) / 16, ntp = 16 * ntb, rows = (int64_t)RB * 16;
        const int64_t kf = (int64_t)t.kpad, kb = (int64_t)t.ns * kTfKsB * 16;   // complex K per direction
        // spectra (FP16 complex, 4 B) once, input spectra (FP32 complex) once, results written
        bytes[0] += t.nbin * (rows * kf * 4 + (int64_t)t.ntile * kf * 8 + t.tsplit_f * ntp * rows * 8);
        bytes[1] += t.nbin * (rows * kb * 4 + (int64_t)t.ntile * kTfKsB * 16 * 8 + (int64_t)t.ns * ntp * rows * 8);
        // per complex MAC eight real FP16 products (re and im of the result, two terms each, hi and
        // lo), 2 flop each.  The four waves issue 16 row blocks (one repeats when RB = 15).
        const double mf = (double)t.nbin * 16 * 16 * ntp * 16.0;
        flops[0] += mf * kf;
        flops[1] += mf * kb;
    }
}

// Transform geometry of group t: nslot planes from slot s0 on.
static TfGeom tf_geom(const sid_psf *p, const FftState &f, const TfGroup &t)
{
    const SensorPlan &sp = p->plan;
    TfGeom g;
    g.n2 = (int)(p->n * p->n);
    g.chp = f.chp;
    g.Lx = (int)sp.Lx;
    g.Ly = (int)sp.Ly;
    g.B = t.B;
    g.nty = t.nty;
    g.ntile = t.ntile;
    g.ntp = (t.ntile + 15) / 16 * 16;
    g.nslot = t.ns;
    g.ngrp = f.ngrp;
    g.nz_in = (int)sp.nz;
    g.pad = (int)p->T;
    g.S = (int)p->S;
    g.XP = sp.XP;
    g.YP = sp.YP;
    g.off = 0;
    g.kslot = f.kslot;
    g.kdim = t.kpad;
    return g;
}

template <int T>
static int tf_in(const float *src, const int *slot_plane, const float *qscale, const TfGeom &g, float2 *out,
                 float *vmax, hipStream_t st)
{
    const size_t lds = sizeof(float2) * kTfCG * tf_lds_cs(T * (T / 2 + 1));
    hipLaunchKernelGGL(k_tf_in<T>, dim3(xcd_grid((int64_t)g.nslot * g.ntile * g.ngrp)), dim3(tf_threads<T>()), lds, st,
                       src, slot_plane, qscale, g, out, vmax);
    SID_TRY_HIP(hipGetLastError());
    return SID_OK;
}

template <int T>
static int tf_out(const float2 *spec, int nsum, int64_t pstride, const int *slot_plane, const float *qscale,
                  const TfGeom &g, float scale, int acc, float *dst, const TfTail &tl, hipStream_t st)
{
    const size_t lds = sizeof(float2) * kTfCG * tf_lds_cs(g.B * (T / 2 + 1));
    const dim3 grid(xcd_grid((int64_t)g.nslot * g.ntile * g.ngrp)), block(tf_threads<T>());
    if (nsum > 1)
        hipLaunchKernelGGL((k_tf_out<T, true>), grid, block, lds, st, spec, nsum, pstride, slot_plane, qscale, g, scale,
                           acc, dst, tl);
    else
        hipLaunchKernelGGL((k_tf_out<T, false>), grid, block, lds, st, spec, nsum, pstride, slot_plane, qscale, g, scale,
                           acc, dst, tl);
    SID_TRY_HIP(hipGetLastError());
    return SID_OK;
}

#define SID_TF_SWITCH(T_, CALL)                          \
    switch (T_) {                                        \
    case 16: return CALL(16);                            \
    case 24: return CALL(24);                            \
    case 32: return CALL(32);                            \
    case 36: return CALL(36);                            \
    case 40: return CALL(40);                            \
    case 48: return CALL(48);                            \
    case 64: return CALL(64);                            \
    default: set_error("internal: tiled length %d", T_); \
        return SID_ERR_ARG;                              \
    }

static int tf_in_any(int T, const float *src, const int *slot_plane, const float *qscale, const TfGeom &g,
                     float2 *out, float *vmax, hipStream_t st)
{
#define SID_TF_IN(L) tf_in<L>(src, slot_plane, qscale, g, out, vmax, st)
    SID_TF_SWITCH(T, SID_TF_IN)
#undef SID_TF_IN
}

static int tf_out_any(int T, const float2 *spec, int nsum, int64_t pstride, const int *slot_plane,
                      const float *qscale, const TfGeom &g, float scale, int acc, float *dst, const TfTail &tl,
                      hipStream_t st)
{
#define SID_TF_OUT(L) tf_out<L>(spec, nsum, pstride, slot_plane, qscale, g, scale, acc, dst, tl, st)
    SID_TF_SWITCH(T, SID_TF_OUT)
#undef SID_TF_OUT
}

template <bool BWD>
static int tf_gemm(const FftState &f, const TfGroup &t, const void *A, const float2 *Bm, int nslot, int KS, int kB,
                   float2 *out, hipStream_t st)
{
    const int RB = f.chp / 16, ntb = (t.ntile + 15) / 16, ntp = ntb * 16;
    const int nsplit = BWD ? t.tsplit_b : t.tsplit_f, chunk = BWD ? t.tchunk_b : t.tchunk_f;
    const int64_t pstride = (int64_t)t.nbin * ntp * f.chp;
    const dim3 grid(xcd_grid((int64_t)t.nbin * nsplit));
    const tu32x4 *a = (const tu32x4 *)A;
#define SID_TF_GEMM(NT) \
    hipLaunchKernelGGL((k_tf_gemm<NT, BWD>), grid, dim3(256), 0, st, a, Bm, t.tmax, t.nbin, nslot, t.ntile, ntp, KS, kB, \
                       RB, f.chp, nsplit, chunk, pstride, out)
    switch (ntb) {
    case 1: SID_TF_GEMM(1); break;
    case 2: SID_TF_GEMM(2); break;
    case 3: SID_TF_GEMM(3); break;
    default: SID_TF_GEMM(4); break;
    }
#undef SID_TF_GEMM
    SID_TRY_HIP(hipGetLastError());
    return SID_OK;
}

static int tf_tmax(const TfGroup &t, int rowlen, hipStream_t st)
{
    hipLaunchKernelGGL(k_tf_tmax, dim3((unsigned)((t.ntile + 3) / 4)), dim3(256), 0, st, t.vmax, t.ntile, rowlen,
                       t.tmax);
    SID_TRY_HIP(hipGetLastError());
    return SID_OK;
}

// Forward image of the FFT planes into f.tout (and f.tout2 for chain 1): in
// each chain group by group, the first writing it and the others adding to
// it (a fixed order); the forward reduction adds the two images.
int tf_forward(sid_psf *p, hipStream_t st)
{
    SensorPlan &sp = p->plan;
    FftState &f = sp.fft;
    const int n2 = (int)(p->n * p->n);
    if (f.nchain == 2) {
        SID_TRY_HIP(hipEventRecord(f.tev_fork, st));
        SID_TRY_HIP(hipStreamWaitEvent(p->aux, f.tev_fork, 0));
    }
    int rc = SID_OK;
    bool first[2] = {true, true};
    for (size_t gi = 0; gi < f.tg.size() && rc == SID_OK; ++gi) {
        const TfGroup &t = f.tg[gi];
        const hipStream_t s = t.chain ? p->aux : st;
        float2 *yout = t.chain ? f.yout2 : f.yout;
        float *tout = t.chain ? f.tout2 : f.tout;
        TfGeom g = tf_geom(p, f, t);
        g.off = -t.Dhi;
        rc = tf_in_any(t.T, sp.vol, f.d_slot_plane + t.s0, f.qscale + (int64_t)t.s0 * n2, g, t.vin, t.vmax, s);
        if (rc == SID_OK) rc = tf_tmax(t, t.ns * f.ngrp, s);
        if (rc != SID_OK) break;
        fft_mark(f, s, 0, true);
        rc = tf_gemm<false>(f, t, t.G, t.vin, 1, t.kpad / 16, t.kpad, yout, s);
        fft_mark(f, s, 0, false);
        if (rc != SID_OK) break;
        TfGeom go = g;
        go.nslot = 1;
        go.off = t.K - 1;
        go.pad = 0;
        go.S = (int)p->NP;
        go.XP = sp.Lx;
        go.YP = sp.Ly;
        rc = tf_out_any(t.T, yout, t.tsplit_f, (int64_t)t.nbin * go.ntp * f.chp, nullptr, nullptr, go,
                        1.0f / ((float)t.T * (float)t.T), first[t.chain] ? 0 : 1, tout, TfTail(), s);
        first[t.chain] = false;
    }
    if (f.nchain == 2) {
        // join even after a failed launch, so that st never runs ahead of work left on the second stream
        SID_TRY_HIP(hipEventRecord(f.tev_join, p->aux));
        SID_TRY_HIP(hipStreamWaitEvent(st, f.tev_join, 0));
    }
    return rc;
}

// Back-projection of the FFT planes into f.dfft (each group its own planes).
// With a tail and every plane of H on this path, the inverse transforms apply
// the backward reduction themselves (*tail_done is set; no dfft round trip).
int tf_backward(sid_psf *p, hipStream_t st, const BwdTail *tail, bool *tail_done)
{
    SensorPlan &sp = p->plan;
    FftState &f = sp.fft;
    const int n2 = (int)(p->n * p->n);
    TfTail tl;
    if (tail && f.n_list_fft == p->nzH && std::getenv("SID_TF_NOFUSE") == nullptr) {
        tl.mode = tail->rl_update ? 2 : 1;
        tl.X = (int)sp.X;
        tl.Y = (int)sp.Y;
        tl.n = (int)p->n;
        tl.htf = sp.htf;
        tl.vol = sp.vol;
        tl.B_col = tail->B_col;
        tl.nf = p->nf;
        tl.check = tail->check;
        tl.mark = tail->mark;
        if (tl.mode == 1 && !tl.B_col) tl.mode = 0;
    }
    if (tail_done) *tail_done = tl.mode != 0;
    if (f.nchain == 2) {
        SID_TRY_HIP(hipEventRecord(f.tev_fork, st));
        SID_TRY_HIP(hipStreamWaitEvent(p->aux, f.tev_fork, 0));
    }
    int rc = SID_OK;
    for (size_t gi = 0; gi < f.tg.size() && rc == SID_OK; ++gi) {
        const TfGroup &t = f.tg[gi];
        const hipStream_t s = t.chain ? p->aux : st;
        float2 *bout = t.chain ? f.bout2 : f.bout;
        TfGeom gi_ = tf_geom(p, f, t);
        gi_.nslot = 1;
        gi_.nz_in = 1;
        gi_.off = t.Dlo;
        gi_.kslot = 0;
        gi_.kdim = kTfKsB * 16;
        rc = tf_in_any(t.T, sp.img, nullptr, nullptr, gi_, t.zin, t.vmax, s);
        if (rc == SID_OK) rc = tf_tmax(t, f.ngrp, s);
        if (rc != SID_OK) break;
        fft_mark(f, s, 1, true);
        rc = tf_gemm<true>(f, t, t.Gb, t.zin, t.ns, kTfKsB, kTfKsB * 16, bout, s);
        fft_mark(f, s, 1, false);
        if (rc != SID_OK) break;
        const TfGeom go = tf_geom(p, f, t);
        rc = tf_out_any(t.T, bout, 1, 0, f.d_slot_plane + t.s0, f.qscale + (int64_t)t.s0 * n2, go,
                        1.0f / ((float)t.T * (float)t.T), 0, f.dfft, tl, s);
    }
    if (f.nchain == 2) {
        SID_TRY_HIP(hipEventRecord(f.tev_join, p->aux));
        SID_TRY_HIP(hipStreamWaitEvent(st, f.tev_join, 0));
    }
    return rc;
}

}  // namespace sid
