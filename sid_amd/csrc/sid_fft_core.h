// Small FFT building blocks shared by the FFT path kernels (sid_fft.hip,
// sid_tfft.hip): complex product, LDS-only barrier, exact small-quotient
// division, and the radix-2/3/4/5 butterflies.
#pragma once
#include <hip/hip_runtime.h>

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

}  // namespace sid
