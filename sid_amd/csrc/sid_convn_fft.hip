// convn(V, k, 'same') by transforms, at every precision where the lengths fit:
// the neuron-kernel wrapping of reconstruct_S.m:122-128,184-190 (forwardFUN =
// F(convn(X, k, 'same')), backwardFUN = convn(B(y), k, 'same')) with the
// 'same' window of MATLAB's convn:
//   out[x, y, z] = sum_{i,j,l} V[x + ox - i, y + oy - j, z + oz - l] k[i, j, l],
//   (ox, oy, oz) = floor(size(k) / 2), V = 0 outside the volume.
//
// The x and y sums are circular convolutions of length Nx >= X + ox and
// Ny >= Y + oy (no wrapped term reaches the 'same' window), taken as products
// of 2-D DFTs; the kernel sits on the grid shifted by (-ox, -oy) (tap (i, j)
// at ((i - ox) mod Nx, (j - oy) mod Ny)), so the window starts at 0 and the
// output rows are aligned.  The z sum (kz <= 8 taps) is done directly in the
// frequency domain before the inverse transforms:
//   O_z(fx, fy) = sum_l K_l(fx, fy) V^_{z + oz - l}(fx, fy).
// Everything is FP32 (rel-L2 error about 1e-7 against an exact convolution:
// below the FP32 direct sum's rounding, where the BF16-split MFMA form has the
// kernel's 2^-9 or 2^-17 quantisation).
//
// Every transform is a "group FFT": N = N1 N2 points held by N2 lanes of a
// wave, N1 per lane (element g + N2 m in lane g), an N1-point register FFT,
// twiddles, one transpose through the wave's own LDS (no workgroup barrier)
// and N1 / N2 register FFTs of N2 points per lane.  Five passes over HBM:
//  1. k_cf_xfwd: two volume rows as one complex sequence (row y + i row y + 1),
//     x-transform, split into the rows' half spectra (bins 0 .. Nx/2), stored
//     by column through LDS: s1[z][fx][y] (16 rows per block: 128-byte runs).
//  2. k_cf_ycol<-1>: y-transform of each column of s1, in place.
//  3. k_cf_zsum: the z sum against the kernel spectra, per frequency, into s2.
//  4. k_cf_ycol<+1>: inverse y-transform of each column of s2, in place,
//     keeping the 'same' rows.
//  5. k_cf_xinv: the Hermitian-completed spectra of two output rows as one
//     complex sequence, inverse x-transform, the 'same' window written out.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "sid_fft_core.h"
#include "sid_internal.h"

namespace sid {

struct CfDims {
    int X, Y, Z, kx, ky, kz, ox, oy, oz;
    int Nx, Ny, Hx, Yp, Pc;
};

constexpr int kCfPairs = 8;   // row pairs per block of the x passes (16 rows)

// LDS order of the wave's own transpose: waits for its LDS operations; the
// memory clobber keeps the compiler from moving LDS accesses across it.
__device__ __forceinline__ void wave_lds_fence() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// DFT of N = N1 N2 points over a group of N2 lanes (lane g holds element
// g + N2 m in x[m]); SGN = -1 forward, +1 inverse (unscaled).  tw[t] =
// exp(-2 pi i t / N) (LDS), tr: the group's N1 x (N2 + 1) LDS transpose area.
// On return lane h holds element h + N2 s + N1 k2 in x[s N2 + k2].
//   X[k1 + N1 k2] = sum_g w_N2^(g k2) [w_N^(g k1) sum_m x_(g + N2 m) w_N1^(m k1)]
template <int N1, int N2, int SGN>
__device__ __forceinline__ void grp_fft(float2 (&x)[N1], int g, float2 *tr, const float2 *tw)
{
    static_assert(N1 % N2 == 0, "group FFT: N2 divides N1");
    constexpr int P = N2 + 1, S = N1 / N2;
    reg_fft<N1, SGN>(x);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) {
        float2 w = tw[g * k1];   // g k1 < N
        if (SGN > 0) w.y = -w.y;
        x[k1] = cmul(x[k1], w);
    }
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) tr[k1 * P + g] = x[k1];
    wave_lds_fence();
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
        for (int j = 0; j < N2; ++j) x[s * N2 + j] = tr[(g + N2 * s) * P + j];
    wave_lds_fence();
#pragma unroll
    for (int s = 0; s < S; ++s) {
        float2 y[N2];
#pragma unroll
        for (int j = 0; j < N2; ++j) y[j] = x[s * N2 + j];
        reg_fft<N2, SGN>(y);
#pragma unroll
        for (int j = 0; j < N2; ++j) x[s * N2 + j] = y[j];
    }
}

__device__ __forceinline__ void cf_load_tw(float2 *tw, const float2 *__restrict__ src, int n)
{
    for (int t = threadIdx.x; t < n; t += blockDim.x) tw[t] = src[t];
}

// LDS of the x passes: twiddles, then one area used first for the groups'
// transposes and then for the row pairs' spectra.  A 64-bit LDS access serves
// 16 lanes per cycle over 16 eight-byte slots; the 16 lanes of the split loop
// are kCfPairs pairs x 16 / kCfPairs bins, so a pitch of N + 16 / kCfPairs
// (N a multiple of 16) puts them on distinct slots.
template <int N1, int N2>
constexpr int cf_x_pitch()
{
    return N1 * N2 + 16 / kCfPairs;
}
template <int N1, int N2>
constexpr size_t cf_x_lds()
{
    constexpr size_t N = (size_t)N1 * N2;
    return sizeof(float2) *
           (N + std::max((size_t)kCfPairs * cf_x_pitch<N1, N2>(), (size_t)kCfPairs * N1 * (N2 + 1)));
}

// Pass 1.  Block (z, 16 rows from y0); group p = row pair (y0 + 2 p, y0 + 2 p + 1).
template <int N1, int N2>
__global__ __launch_bounds__(kCfPairs * N2) void k_cf_xfwd(const float *__restrict__ V, CfDims g,
                                                           const float2 *__restrict__ twx, float2 *__restrict__ s1)
{
    constexpr int N = N1 * N2, SP = cf_x_pitch<N1, N2>();
    extern __shared__ float2 cfs[];
    float2 *tw = cfs, *area = cfs + N;
    const int nyb = (g.Y + 2 * kCfPairs - 1) / (2 * kCfPairs);
    const int z = (int)blockIdx.x / nyb, y0 = ((int)blockIdx.x - z * nyb) * 2 * kCfPairs;
    const int p = (int)threadIdx.x / N2, lg = (int)threadIdx.x % N2, y = y0 + 2 * p;
    const float *r0 = V + ((int64_t)z * g.Y + y) * g.X, *r1 = r0 + g.X;
    float2 x[N1];
#pragma unroll
    for (int m = 0; m < N1; ++m) {
        const int c = lg + N2 * m;
        x[m].x = (c < g.X && y < g.Y) ? r0[c] : 0.0f;
        x[m].y = (c < g.X && y + 1 < g.Y) ? r1[c] : 0.0f;
    }
    cf_load_tw(tw, twx, N);
    __syncthreads();
    grp_fft<N1, N2, -1>(x, lg, area + p * N1 * (N2 + 1), tw);
    __syncthreads();   // every group's transpose done: the area now holds spectra
#pragma unroll
    for (int s = 0; s < N1 / N2; ++s)
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) area[p * SP + lg + N2 * s + N1 * k2] = x[s * N2 + k2];
    __syncthreads();
    // Z = A + i B with A, B the rows' Hermitian spectra:
    //   A(k) = (Z(k) + conj Z(-k)) / 2,  B(k) = (Z(k) - conj Z(-k)) / (2 i)
    constexpr int NE = kCfPairs * (N / 2 + 1), NT = kCfPairs * N2, IT = (NE + NT - 1) / NT;
    float2 *dst = s1 + (int64_t)z * g.Hx * g.Pc + y0;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int e = (int)threadIdx.x + it * NT;
        if (e < NE) {
            const int fx = e / kCfPairs, q = e % kCfPairs;
            const float2 z1 = area[q * SP + fx];
            const float2 z2 = area[q * SP + (fx == 0 ? 0 : N - fx)];
            const float4 o = make_float4(0.5f * (z1.x + z2.x), 0.5f * (z1.y - z2.y), 0.5f * (z1.y + z2.y),
                                         0.5f * (z2.x - z1.x));
            *reinterpret_cast<float4 *>(dst + (int64_t)fx * g.Pc + 2 * q) = o;
        }
    }
}

// Groups per block of the column passes (LDS of the transposes within ~80 KB).
template <int N1, int N2>
constexpr int cf_col_groups()
{
    return (size_t)N1 * (N2 + 1) * sizeof(float2) * (256 / N2) <= 80 * 1024 ? 256 / N2 : 128 / N2;
}
template <int N1, int N2>
constexpr size_t cf_col_lds()
{
    return sizeof(float2) * ((size_t)N1 * N2 + (size_t)cf_col_groups<N1, N2>() * N1 * (N2 + 1));
}

// Passes 2 and 4: one column (z, fx) per group, in place.  Forward: rows
// y < Y in (the rest zero), all Ny bins out.  Inverse: all bins in, rows y < Y
// out and zeros up to Yp (rows the x pass reads).
template <int N1, int N2, int SGN>
__global__ __launch_bounds__((cf_col_groups<N1, N2>() * N2)) void k_cf_ycol(CfDims g,
                                                                           const float2 *__restrict__ twy,
                                                                           float2 *__restrict__ s)
{
    constexpr int N = N1 * N2, GB = cf_col_groups<N1, N2>();
    extern __shared__ float2 cfs[];
    float2 *tw = cfs;
    const int q = (int)threadIdx.x / N2, lg = (int)threadIdx.x % N2;
    const int64_t col = (int64_t)blockIdx.x * GB + q;
    const bool live = col < (int64_t)g.Z * g.Hx;
    float2 *c = s + col * g.Pc;
    const int nin = SGN < 0 ? g.Y : N;
    float2 x[N1];
#pragma unroll
    for (int m = 0; m < N1; ++m) {
        const int r = lg + N2 * m;
        x[m] = (live && r < nin) ? c[r] : make_float2(0.0f, 0.0f);
    }
    cf_load_tw(tw, twy, N);
    __syncthreads();
    if (!live) return;
    grp_fft<N1, N2, SGN>(x, lg, cfs + N + q * N1 * (N2 + 1), tw);
#pragma unroll
    for (int s = 0; s < N1 / N2; ++s)
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) {
            const int k = lg + N2 * s + N1 * k2;
            if (SGN < 0 || k < g.Y) c[k] = x[s * N2 + k2];
        }
    if (SGN > 0 && lg < g.Yp - g.Y) c[g.Y + lg] = make_float2(0.0f, 0.0f);
}

// Pass 3: the z sum per frequency, s2[zo] = sum_l K_l s1[zo + oz - l].  A
// thread owns two neighbouring y bins (16 bytes) of one column fx and kCfZc
// output planes: the kCfZc + kz - 1 input values it needs are loaded first
// (all in flight together), with the kernel spectra reversed, Kr[l'] =
// K_{kz - 1 - l'}, so that out[j] = sum_l' Kr[l'] in[j + l'] has constant
// register indices.  Items (column pairs block, z chunk) of one XCD are
// contiguous, so the kz - 1 planes two chunks share are read once from HBM.
constexpr int kCfZc = 8;
template <int KZM>
__global__ __launch_bounds__(256) void k_cf_zsum(CfDims g, const float2 *__restrict__ s1,
                                                 const float2 *__restrict__ kf, float2 *__restrict__ s2)
{
    const int npair = g.Ny / 2, ncb = (g.Hx * npair + 255) / 256, nch = (g.Z + kCfZc - 1) / kCfZc;
    const int item = xcd_item(ncb * nch);
    if (item < 0) return;
    const int cb = item / nch, ch = item - cb * nch;
    const int e = cb * 256 + (int)threadIdx.x;
    if (e >= g.Hx * npair) return;
    const int fx = e / npair, fy = 2 * (e - fx * npair);
    float4 kr[KZM];
#pragma unroll
    for (int l = 0; l < KZM; ++l)
        kr[l] = l < g.kz ? *reinterpret_cast<const float4 *>(kf + ((int64_t)(g.kz - 1 - l) * g.Hx + fx) * g.Ny + fy)
                         : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const int zlo = ch * kCfZc, zb = zlo + g.oz - (g.kz - 1);
    const int64_t zs = (int64_t)g.Hx * g.Pc;
    const float2 *src = s1 + (int64_t)fx * g.Pc + fy;
    float4 in[kCfZc + KZM - 1];
#pragma unroll
    for (int t = 0; t < kCfZc + KZM - 1; ++t) {
        const int zi = zb + t;
        in[t] = (zi >= 0 && zi < g.Z && t < kCfZc + g.kz - 1) ? *reinterpret_cast<const float4 *>(src + zi * zs)
                                                             : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    float2 *dst = s2 + (int64_t)fx * g.Pc + fy;
#pragma unroll
    for (int j = 0; j < kCfZc; ++j) {
        if (zlo + j >= g.Z) break;
        float4 o = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int l = 0; l < KZM; ++l) {
            if (l >= g.kz) break;   // (a zero weight times a non-finite input would still be NaN)
            const float4 k = kr[l], v = in[j + l];
            o.x += k.x * v.x - k.y * v.y;
            o.y += k.x * v.y + k.y * v.x;
            o.z += k.z * v.z - k.w * v.w;
            o.w += k.z * v.w + k.w * v.z;
        }
        *reinterpret_cast<float4 *>(dst + (zlo + j) * zs) = o;
    }
}

// Pass 5.  Block (z, 16 rows from y0): the half spectra A, B of rows y0 + 2 p
// and y0 + 2 p + 1 give Z(k) = A(k) + i B(k), Z(Nx - k) = conj A(k) + i conj B(k);
// the inverse transform is row y0 + 2 p + i row y0 + 2 p + 1.
template <int N1, int N2>
__global__ __launch_bounds__(kCfPairs * N2) void k_cf_xinv(CfDims g, const float2 *__restrict__ s2,
                                                           const float2 *__restrict__ twx, float *__restrict__ out)
{
    constexpr int N = N1 * N2, SP = cf_x_pitch<N1, N2>();
    extern __shared__ float2 cfs[];
    float2 *tw = cfs, *area = cfs + N;
    const int nyb = (g.Y + 2 * kCfPairs - 1) / (2 * kCfPairs);
    const int z = (int)blockIdx.x / nyb, y0 = ((int)blockIdx.x - z * nyb) * 2 * kCfPairs;
    // all of the block's loads first (Hx = N / 2 + 1 bins x 8 pairs), then the LDS writes
    constexpr int NE = kCfPairs * (N / 2 + 1), NT = kCfPairs * N2, IT = (NE + NT - 1) / NT;
    const float2 *src = s2 + (int64_t)z * g.Hx * g.Pc + y0;
    float4 v[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int e = (int)threadIdx.x + it * NT;
        if (e < NE) v[it] = *reinterpret_cast<const float4 *>(src + (int64_t)(e / kCfPairs) * g.Pc + 2 * (e % kCfPairs));
    }
    cf_load_tw(tw, twx, N);
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int e = (int)threadIdx.x + it * NT;
        if (e < NE) {
            const int fx = e / kCfPairs, q = e % kCfPairs;
            area[q * SP + fx] = make_float2(v[it].x - v[it].w, v[it].y + v[it].z);
            if (fx > 0 && fx < N - fx) area[q * SP + N - fx] = make_float2(v[it].x + v[it].w, v[it].z - v[it].y);
        }
    }
    __syncthreads();
    const int p = (int)threadIdx.x / N2, lg = (int)threadIdx.x % N2, y = y0 + 2 * p;
    float2 x[N1];
#pragma unroll
    for (int m = 0; m < N1; ++m) x[m] = area[p * SP + lg + N2 * m];
    __syncthreads();   // the area turns into the groups' transposes
    grp_fft<N1, N2, 1>(x, lg, area + p * N1 * (N2 + 1), tw);
    float *r0 = out + ((int64_t)z * g.Y + y) * g.X, *r1 = r0 + g.X;
#pragma unroll
    for (int s = 0; s < N1 / N2; ++s)
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) {
            const int xo = lg + N2 * s + N1 * k2;
            if (xo < g.X) {
                if (y < g.Y) r0[xo] = x[s * N2 + k2].x;
                if (y + 1 < g.Y) r1[xo] = x[s * N2 + k2].y;
            }
        }
}

// Kernel spectra, tap (i, j, l) at ((i - ox) mod Nx, (j - oy) mod Ny) of
// plane l: x transforms of the kernel rows here (row j at grid row
// r = (j - oy) mod Ny, other rows zero), w_Nx = exp(-2 pi i / Nx) from the x
// transform's twiddle table, scaled by 1 / (Nx Ny) for the two unnormalised
// inverse transforms; then the forward column pass over kf[l][fx][Ny].
__global__ void k_cf_kgrid(const float *__restrict__ k, CfDims g, const float2 *__restrict__ twx, float scale,
                           float2 *__restrict__ kf)
{
    const int64_t n = (int64_t)g.kz * g.Hx * g.Ny;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(e % g.Ny);
        const int64_t lf = e / g.Ny;
        const int fx = (int)(lf % g.Hx), l = (int)(lf / g.Hx);
        const int j = (r + g.oy) % g.Ny;
        float2 s = make_float2(0.0f, 0.0f);
        if (j < g.ky) {
            const float *row = k + ((int64_t)l * g.ky + j) * g.kx;
            int t = (int)(((int64_t)fx * (g.Nx - g.ox)) % g.Nx);   // fx (i - ox) mod Nx at i = 0
            for (int i = 0; i < g.kx; ++i) {
                const float2 w = twx[t];
                s.x += row[i] * w.x;
                s.y += row[i] * w.y;
                t += fx;
                if (t >= g.Nx) t -= g.Nx;
            }
        }
        kf[e] = make_float2(s.x * scale, s.y * scale);
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

// Transform lengths of the group FFT: N = N1 N2 with N2 | N1 <= 64.
static const int kCfLengths[] = {128, 256, 512, 768, 1024, 2048};

static int cf_length(int n)
{
    for (int L : kCfLengths)
        if (L >= n) return L;
    return 0;
}

static bool cf_shape(const int64_t vdims[3], const int kd[3], int &Nx, int &Ny)
{
    if (vdims[0] < 1 || vdims[1] < 1 || vdims[2] < 1 || vdims[0] > 4096 || vdims[1] > 4096 || vdims[2] > 65535)
        return false;
    if (kd[0] < 1 || kd[1] < 1 || kd[2] < 1 || kd[2] > 8) return false;
    const int X = (int)vdims[0], Y = (int)vdims[1];
    Nx = cf_length(std::max(X + kd[0] / 2, kd[0]));
    Ny = cf_length(std::max(Y + kd[1] / 2, kd[1]));
    return Nx > 0 && Ny > 0;
}

bool convn_fft_fits(const int64_t vdims[3], const int kd[3])
{
    int nx, ny;
    return cf_shape(vdims, kd, nx, ny);
}

bool convn_fft_enabled()
{
    const char *e = std::getenv("SID_CONVN_FFT");
    return !(e && e[0] == '0');
}

void convn_fft_free(ConvnFft &c)
{
    fft_plan_free(c.px);
    fft_plan_free(c.py);
    if (c.kf) (void)hipFree(c.kf);
    if (c.s1) (void)hipFree(c.s1);
    if (c.s2) (void)hipFree(c.s2);
    c = ConvnFft();
}

static CfDims cf_dims(const ConvnFft &c)
{
    CfDims g;
    g.X = c.X;
    g.Y = c.Y;
    g.Z = c.Z;
    g.kx = c.kx;
    g.ky = c.ky;
    g.kz = c.kz;
    g.ox = c.kx / 2;
    g.oy = c.ky / 2;
    g.oz = c.kz / 2;
    g.Nx = c.Nx;
    g.Ny = c.Ny;
    g.Hx = c.Hx;
    g.Yp = c.Yp;
    g.Pc = c.Pc;
    return g;
}

static int cf_y(const ConvnFft &c, const CfDims &g, bool fwd, float2 *buf, hipStream_t st);

int convn_fft_prepare(ConvnFft &c, const int64_t vdims[3], const float *k_dev, const int kd[3], hipStream_t st)
{
    int Nx = 0, Ny = 0;
    if (!cf_shape(vdims, kd, Nx, Ny)) {
        set_error("convn: volume %lld x %lld x %lld with kernel %d x %d x %d does not fit the transform form",
                  (long long)vdims[0], (long long)vdims[1], (long long)vdims[2], kd[0], kd[1], kd[2]);
        return SID_ERR_DIMS;
    }
    const int X = (int)vdims[0], Y = (int)vdims[1], Z = (int)vdims[2];
    const bool same = c.X == X && c.Y == Y && c.Z == Z && c.kx == kd[0] && c.ky == kd[1] && c.kz == kd[2];
    if (!same) {
        if (c.Nx != Nx) {
            fft_plan_free(c.px);
            SID_TRY(fft_plan_make(c.px, Nx));
        }
        if (c.Ny != Ny) {
            fft_plan_free(c.py);
            SID_TRY(fft_plan_make(c.py, Ny));
        }
        const int Hx = Nx / 2 + 1;
        const size_t nkf = (size_t)Hx * kd[2] * Ny;
        const size_t okf = c.kf ? (size_t)c.Hx * c.kz * c.Ny : 0;
        if (nkf != okf) {
            if (c.kf) (void)hipFree(c.kf);
            c.kf = nullptr;
            SID_TRY_HIP(hipMalloc(&c.kf, sizeof(float2) * nkf));
        }
        const int Yp = (Y + 2 * kCfPairs - 1) / (2 * kCfPairs) * (2 * kCfPairs);
        const int Pc = std::max(Ny, Yp);
        const size_t ns = (size_t)Z * Hx * Pc;
        if (ns > c.s_elems) {
            if (c.s1) (void)hipFree(c.s1);
            if (c.s2) (void)hipFree(c.s2);
            c.s1 = c.s2 = nullptr;
            c.s_elems = 0;
            SID_TRY_HIP(hipMalloc(&c.s1, sizeof(float2) * ns));
            SID_TRY_HIP(hipMalloc(&c.s2, sizeof(float2) * ns));
            c.s_elems = ns;
        }
        c.X = X;
        c.Y = Y;
        c.Z = Z;
        c.kx = kd[0];
        c.ky = kd[1];
        c.kz = kd[2];
        c.Nx = Nx;
        c.Ny = Ny;
        c.Hx = Hx;
        c.Yp = Yp;
        c.Pc = Pc;
        c.kf_ready = false;
    }
    if (!c.kf_ready) {
        const CfDims g = cf_dims(c);
        const int64_t nkf = (int64_t)c.Hx * c.kz * c.Ny;
        hipLaunchKernelGGL(k_cf_kgrid, dim3((unsigned)std::min<int64_t>(4096, (nkf + 255) / 256)), dim3(256), 0, st,
                           k_dev, g, c.px.tw_fwd, (float)(1.0 / ((double)c.Nx * (double)c.Ny)), c.kf);
        SID_TRY_HIP(hipGetLastError());
        CfDims gk = g;   // the kernel planes as kz planes of Ny-row columns
        gk.Y = gk.Yp = gk.Pc = c.Ny;
        gk.Z = c.kz;
        SID_TRY(cf_y(c, gk, true, c.kf, st));
        c.kf_ready = true;
    }
    return SID_OK;
}

template <int N1, int N2>
static int cf_launch_x(const ConvnFft &c, const CfDims &g, bool fwd, const float *in, float *out, hipStream_t st)
{
    const size_t lds = cf_x_lds<N1, N2>();
    const unsigned nblk = (unsigned)((int64_t)c.Z * ((c.Y + 2 * kCfPairs - 1) / (2 * kCfPairs)));
    if (fwd) {
        SID_TRY_HIP(hipFuncSetAttribute((const void *)k_cf_xfwd<N1, N2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)lds));
        hipLaunchKernelGGL((k_cf_xfwd<N1, N2>), dim3(nblk), dim3(kCfPairs * N2), lds, st, in, g, c.px.tw_fwd, c.s1);
    } else {
        SID_TRY_HIP(hipFuncSetAttribute((const void *)k_cf_xinv<N1, N2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)lds));
        hipLaunchKernelGGL((k_cf_xinv<N1, N2>), dim3(nblk), dim3(kCfPairs * N2), lds, st, g, c.s2, c.px.tw_fwd, out);
    }
    SID_TRY_HIP(hipGetLastError());
    return SID_OK;
}

template <int N1, int N2>
static int cf_launch_y(const CfDims &g, const float2 *tw, bool fwd, float2 *buf, hipStream_t st)
{
    constexpr int GB = cf_col_groups<N1, N2>();
    const size_t lds = cf_col_lds<N1, N2>();
    const int64_t ncol = (int64_t)g.Z * g.Hx, nblk = (ncol + GB - 1) / GB;
    const void *kern = fwd ? (const void *)k_cf_ycol<N1, N2, -1> : (const void *)k_cf_ycol<N1, N2, 1>;
    SID_TRY_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (fwd) hipLaunchKernelGGL((k_cf_ycol<N1, N2, -1>), dim3((unsigned)nblk), dim3(GB * N2), lds, st, g, tw, buf);
    else hipLaunchKernelGGL((k_cf_ycol<N1, N2, 1>), dim3((unsigned)nblk), dim3(GB * N2), lds, st, g, tw, buf);
    SID_TRY_HIP(hipGetLastError());
    return SID_OK;
}

static int cf_zsum(const ConvnFft &c, const CfDims &g, hipStream_t st)
{
    const int64_t ncb = ((int64_t)c.Hx * (c.Ny / 2) + 255) / 256, nch = (c.Z + kCfZc - 1) / kCfZc;
    const dim3 grid(xcd_grid(ncb * nch));
    if (c.kz <= 2) hipLaunchKernelGGL(k_cf_zsum<2>, grid, dim3(256), 0, st, g, c.s1, c.kf, c.s2);
    else if (c.kz <= 4) hipLaunchKernelGGL(k_cf_zsum<4>, grid, dim3(256), 0, st, g, c.s1, c.kf, c.s2);
    else hipLaunchKernelGGL(k_cf_zsum<8>, grid, dim3(256), 0, st, g, c.s1, c.kf, c.s2);
    SID_TRY_HIP(hipGetLastError());
    return SID_OK;
}

static int cf_x(const ConvnFft &c, const CfDims &g, bool fwd, const float *in, float *out, hipStream_t st)
{
    switch (c.Nx) {
    case 128: return cf_launch_x<16, 8>(c, g, fwd, in, out, st);
    case 256: return cf_launch_x<16, 16>(c, g, fwd, in, out, st);
    case 512: return cf_launch_x<32, 16>(c, g, fwd, in, out, st);
    case 768: return cf_launch_x<48, 16>(c, g, fwd, in, out, st);
    case 1024: return cf_launch_x<32, 32>(c, g, fwd, in, out, st);
    case 2048: return cf_launch_x<64, 32>(c, g, fwd, in, out, st);
    }
    set_error("convn: transform length %d", c.Nx);
    return SID_ERR_DIMS;
}

// column transforms of g.Z x g.Hx columns of buf (pitch g.Pc), length c.Ny
static int cf_y(const ConvnFft &c, const CfDims &g, bool fwd, float2 *buf, hipStream_t st)
{
    const float2 *tw = c.py.tw_fwd;
    switch (c.Ny) {
    case 128: return cf_launch_y<16, 8>(g, tw, fwd, buf, st);
    case 256: return cf_launch_y<16, 16>(g, tw, fwd, buf, st);
    case 512: return cf_launch_y<32, 16>(g, tw, fwd, buf, st);
    case 768: return cf_launch_y<48, 16>(g, tw, fwd, buf, st);
    case 1024: return cf_launch_y<32, 32>(g, tw, fwd, buf, st);
    case 2048: return cf_launch_y<64, 32>(g, tw, fwd, buf, st);
    }
    set_error("convn: transform length %d", c.Ny);
    return SID_ERR_DIMS;
}

int convn_fft_run(ConvnFft &c, const float *in, float *out, hipStream_t st)
{
    if (!c.kf_ready || !c.s1) {
        set_error("convn: transform form not prepared");
        return SID_ERR_ARG;
    }
    const CfDims g = cf_dims(c);
    SID_TRY(cf_x(c, g, true, in, nullptr, st));
    SID_TRY(cf_y(c, g, true, c.s1, st));
    SID_TRY(cf_zsum(c, g, st));
    SID_TRY(cf_y(c, g, false, c.s2, st));
    SID_TRY(cf_x(c, g, false, nullptr, out, st));
    return SID_OK;
}

// Work arrays of the handle-free entry (sid_convn_same_device), one set per
// device, kept for the process: calls on one device take turns.
static std::mutex g_cf_mutex;
static std::vector<ConvnFft> g_cf_work;

int convn_fft_oneshot(int device, const float *in, const int64_t vdims[3], const float *k_dev, const int kd[3],
                      float *out, hipStream_t st)
{
    std::lock_guard<std::mutex> lock(g_cf_mutex);
    if ((int)g_cf_work.size() <= device) g_cf_work.resize(device + 1);
    ConvnFft &c = g_cf_work[device];
    c.kf_ready = false;   // the kernel may differ from the last call's
    int rc = convn_fft_prepare(c, vdims, k_dev, kd, st);
    if (rc == SID_OK) rc = convn_fft_run(c, in, out, st);
    const hipError_t e = hipStreamSynchronize(st);
    if (rc != SID_OK) return rc;
    if (e != hipSuccess) {
        set_error("convn: %s", hipGetErrorString(e));
        return SID_ERR_HIP;
    }
    return SID_OK;
}

}  // namespace sid
