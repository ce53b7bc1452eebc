// Small FFT building blocks shared by the FFT kernels (sid_fft.hip,
// sid_tfft.hip, sid_convn_fft.hip): complex product, LDS-only barrier, exact
// small-quotient division, the radix-2/3/4/5 butterflies, the Stockham stages
// of sequences held in LDS, and the XCD-contiguous work mapping.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace sid {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

// Workgroup barrier for LDS traffic only: waits for this wave's LDS operations
// and meets the other waves, without waiting for outstanding global loads
// (the register prefetch of the persistent kernels stays in flight).  The
// memory clobber keeps the compiler from moving memory accesses across it.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// a / b for 0 <= a < 2^22, b >= 1, with rb = 1.0f / b: the float quotient is
// off by at most one and corrected exactly (no integer division sequence).
__device__ __forceinline__ int qdiv(int a, int b, float rb)
{
    int q = (int)((float)a * rb);
    q -= (q * b > a);
    q += ((q + 1) * b <= a);
    return q;
}

template <int R>
__device__ __forceinline__ void small_dft(float2 *v, float sgn)
{
    if constexpr (R == 2) {
        const float2 a = v[0], b = v[1];
        v[0] = make_float2(a.x + b.x, a.y + b.y);
        v[1] = make_float2(a.x - b.x, a.y - b.y);
    } else if constexpr (R == 4) {
        const float2 a = v[0], b = v[1], c = v[2], d = v[3];
        const float2 s0 = make_float2(a.x + c.x, a.y + c.y), s1 = make_float2(a.x - c.x, a.y - c.y);
        const float2 s2 = make_float2(b.x + d.x, b.y + d.y), s3 = make_float2(b.x - d.x, b.y - d.y);
        // multiply s3 by (sgn * i)
        const float2 s3i = make_float2(-sgn * s3.y, sgn * s3.x);
        v[0] = make_float2(s0.x + s2.x, s0.y + s2.y);
        v[2] = make_float2(s0.x - s2.x, s0.y - s2.y);
        v[1] = make_float2(s1.x + s3i.x, s1.y + s3i.y);
        v[3] = make_float2(s1.x - s3i.x, s1.y - s3i.y);
    } else if constexpr (R == 3) {
        // w = exp(sgn 2 pi i / 3) = -1/2 + sgn i sqrt(3)/2
        const float c1 = -0.5f, s1 = sgn * 0.86602540378443865f;
        const float2 a = v[0], b = v[1], c = v[2];
        const float2 t = make_float2(b.x + c.x, b.y + c.y), u = make_float2(b.x - c.x, b.y - c.y);
        v[0] = make_float2(a.x + t.x, a.y + t.y);
        const float2 m = make_float2(a.x + c1 * t.x, a.y + c1 * t.y);
        // v1 = m + i s1 u, v2 = m - i s1 u
        v[1] = make_float2(m.x - s1 * u.y, m.y + s1 * u.x);
        v[2] = make_float2(m.x + s1 * u.y, m.y - s1 * u.x);
    } else {
        // R == 5: w^e = cos(2 pi e / 5) + sgn i sin(2 pi e / 5)
        const float c1 = 0.30901699437494742f, c2 = -0.80901699437494742f;
        const float s1 = sgn * 0.95105651629515357f, s2 = sgn * 0.58778525229247313f;
        const float2 a = v[0];
        const float2 t1 = make_float2(v[1].x + v[4].x, v[1].y + v[4].y), u1 = make_float2(v[1].x - v[4].x, v[1].y - v[4].y);
        const float2 t2 = make_float2(v[2].x + v[3].x, v[2].y + v[3].y), u2 = make_float2(v[2].x - v[3].x, v[2].y - v[3].y);
        v[0] = make_float2(a.x + t1.x + t2.x, a.y + t1.y + t2.y);
        const float2 m1 = make_float2(a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y);
        const float2 m2 = make_float2(a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y);
        // X1 = m1 + i (s1 u1 + s2 u2), X4 = m1 - i (...); X2 = m2 + i (s2 u1 - s1 u2), X3 = m2 - i (...)
        const float2 n1 = make_float2(s1 * u1.x + s2 * u2.x, s1 * u1.y + s2 * u2.y);
        const float2 n2 = make_float2(s2 * u1.x - s1 * u2.x, s2 * u1.y - s1 * u2.y);
        v[1] = make_float2(m1.x - n1.y, m1.y + n1.x);
        v[4] = make_float2(m1.x + n1.y, m1.y - n1.x);
        v[2] = make_float2(m2.x - n2.y, m2.y + n2.x);
        v[3] = make_float2(m2.x + n2.y, m2.y - n2.x);
    }
}

// Two adjacent floats at a 4-byte aligned address as one 8-byte access
// (multi-dword global accesses need only dword alignment).
__device__ __forceinline__ float2 ld_f2_a4(const float *p)
{
    float2 v;
    __builtin_memcpy(&v, p, 8);
    return v;
}
__device__ __forceinline__ void st_f2_a4(float *p, float2 v) { __builtin_memcpy(p, &v, 8); }

// Radices of one transform (outermost first) and sequences per block.
struct FftDesc {
    int L, nst, spb;
    int rad[12];
};

// One Stockham stage over nseq sequences of length L in LDS, element t of
// sequence s at s * ss + t * es.  Consecutive threads take consecutive
// sequences when the sequences are interleaved (es > 1), consecutive elements
// otherwise, so LDS accesses stay on distinct banks.
template <int R>
__device__ __forceinline__ void stockham_stage_2d(const float2 *__restrict__ in, float2 *__restrict__ out, int L,
                                                  int Ns, const float2 *__restrict__ tw, float sgn, int nseq, int ss,
                                                  int es)
{
    const int nb = L / R;
    const int tot = nseq * nb;
    const int step = L / (Ns * R);   // step * k * r < L: no wrap of the twiddle index
    const float rns = 1.0f / (float)Ns, rseq = 1.0f / (float)nseq;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        // neighbouring lanes take neighbouring lines: with an odd line pitch (or
        // unit line stride) every read and write of a stage is bank-conflict free
        const int j = qdiv(e, nseq, rseq), sq = e - j * nseq;
        const float2 *src = in + sq * ss;
        float2 *dst = out + sq * ss;
        const int jq = qdiv(j, Ns, rns), k = j - jq * Ns;
        const int tstep = step * k;
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const float2 x = src[(j + r * nb) * es];
            v[r] = r == 0 ? x : cmul(x, tw[tstep * r]);
        }
        small_dft<R>(v, sgn);
        const int base = jq * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[(base + r * Ns) * es] = v[r];
    }
}

// All stages of one direction; returns the buffer holding the result.
__device__ inline float2 *fft_pass_2d(float2 *a, float2 *b, const FftDesc &d, const float2 *__restrict__ tw, float sgn,
                               int nseq, int ss, int es)
{
    int Ns = 1;
    float2 *in = a, *out = b;
    for (int st = 0; st < d.nst; ++st) {
        const int R = d.rad[st];
        if (R == 4) stockham_stage_2d<4>(in, out, d.L, Ns, tw, sgn, nseq, ss, es);
        else if (R == 2) stockham_stage_2d<2>(in, out, d.L, Ns, tw, sgn, nseq, ss, es);
        else if (R == 3) stockham_stage_2d<3>(in, out, d.L, Ns, tw, sgn, nseq, ss, es);
        else stockham_stage_2d<5>(in, out, d.L, Ns, tw, sgn, nseq, ss, es);
        lds_barrier();
        Ns *= R;
        float2 *t = in; in = out; out = t;
    }
    return in;
}

// Blocks are dealt to the 8 XCDs round robin and each XCD has its own L2:
// item i of n goes to XCD i / ceil(n / 8), so neighbouring items (which share
// cache lines) run on one XCD.  Launch xcd_grid(n) blocks; surplus blocks get -1.
__device__ __forceinline__ int xcd_item(int n)
{
    const int per = (n + 7) >> 3;
    const int it = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    return it < n ? it : -1;
}
static inline unsigned xcd_grid(int64_t n) { return (unsigned)(8 * ((n + 7) / 8)); }

// ---------------------------------------------------------------------------
// Register FFT of compile-time length N (2/3/5-smooth): Stockham stages with
// radices 4, then 2, 3, 5.  All indices are compile-time constants after
// unrolling, so the sequence stays in registers, and the twiddles
// exp(sgn 2 pi i t / N) come from a table evaluated by the compiler (a
// 24-term Taylor series in double on [-pi, pi], multiples of pi/2 exact),
// which folds into literal operands: no table loads, no scalar registers.
// ---------------------------------------------------------------------------
constexpr double kCxPi = 3.14159265358979323846264338327950288;
constexpr double cx_taylor(double x, bool sine)
{
    double term = sine ? x : 1.0, sum = term;
    for (int k = 1; k < 24; ++k) {
        term *= -x * x / (sine ? (2.0 * k) * (2.0 * k + 1) : (2.0 * k - 1) * (2.0 * k));
        sum += term;
    }
    return sum;
}
template <int N, int SGN>
struct TwTable {
    float re[N], im[N];
    constexpr TwTable() : re(), im()
    {
        for (int t = 0; t < N; ++t) {
            if ((4 * t) % N == 0) {   // quarter turns: exact
                const int q = 4 * t / N;
                re[t] = q == 0 ? 1.f : q == 2 ? -1.f : 0.f;
                im[t] = q == 1 ? (float)SGN : q == 3 ? (float)-SGN : 0.f;
                continue;
            }
            double x = 2.0 * kCxPi * t / N;
            if (x > kCxPi) x -= 2.0 * kCxPi;
            re[t] = (float)cx_taylor(x, false);
            im[t] = (float)(SGN * cx_taylor(x, true));
        }
    }
};
template <int N, int SGN>
inline constexpr TwTable<N, SGN> kTw{};

template <int N, int Ns, int R, int SGN>
__device__ __forceinline__ void rf_stage(float2 (&x)[N])
{
    float2 y[N];
#pragma unroll
    for (int j = 0; j < N / R; ++j) {
        const int k = j % Ns;
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const float2 a = x[j + r * (N / R)];
            const int t = k * r * (N / (Ns * R));
            v[r] = (r == 0 || k == 0) ? a : cmul(a, make_float2(kTw<N, SGN>.re[t], kTw<N, SGN>.im[t]));
        }
        small_dft<R>(v, (float)SGN);
        const int base = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) y[base + r * Ns] = v[r];
    }
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = y[i];
}

template <int N, int Ns, int Rem, int SGN>
struct RegFft {
    static __device__ __forceinline__ void run(float2 (&x)[N])
    {
        constexpr int R = Rem % 4 == 0 ? 4 : Rem % 2 == 0 ? 2 : Rem % 3 == 0 ? 3 : 5;
        static_assert(Rem % R == 0, "register FFT lengths are 2/3/5-smooth");
        rf_stage<N, Ns, R, SGN>(x);
        RegFft<N, Ns * R, Rem / R, SGN>::run(x);
    }
};
template <int N, int Ns, int SGN>
struct RegFft<N, Ns, 1, SGN> {
    static __device__ __forceinline__ void run(float2 (&)[N]) {}
};
// Unscaled DFT in place: SGN = -1 forward, +1 inverse.
template <int N, int SGN>
__device__ __forceinline__ void reg_fft(float2 (&x)[N])
{
    RegFft<N, 1, N, SGN>::run(x);
}

}  // namespace sid
