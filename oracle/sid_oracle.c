/*
 * ORACLE — test infrastructure only (the checker and the CPU baseline, never
 * the product).  Plain-C restatement of SID's CPU projectors and ISRA loop.
 *
 *   orc_forward_*   external/LFrecon_util/forwardProjectACC.m:14-29 (f64: FP64 PSF;
 *                   h32: FP32 PSF, FP64 arithmetic; f32: the reference's single precision)
 *                   (forwardProjectGPU.m:16-28 when CAindex is NULL: full M)
 *   orc_backward_*  external/LFrecon_util/backwardProjectACC_new.m:14-26
 *                   (backwardProjectGPU_new.m:16-26 when CAindex is NULL)
 *   orc_isra_f64    external/LFrecon_util/deconvRL.m:46-52,85-87 (basic, lambda 0)
 *
 * Layouts are MATLAB column-major: H(i,j,aa,bb,cc) at
 * i + M*(j + M*(aa + n*(bb + n*cc))), V(x,y,c) at x + X*(y + Y*c), O(x,y) at
 * x + X*y, CAindex(c,k) at c + Nz*(k-1) with 1-based values.
 *
 * conv2(t, K, 'same') of the sublattice-masked plane t is evaluated as the
 * same sum restricted to the nonzero entries of t (the masked-out entries
 * contribute exact zeros): O[p] += V[q] * K[p - q + floor(Mc/2)].  The
 * backward uses the 180-degree-rotated kernel, i.e. B[q] = sum_p O[p] *
 * K[p - q + (Mc - 1 - floor(Mc/2))].
 *
 * Parallelisation: OpenMP over output columns (each thread owns a disjoint
 * column range, so the accumulation order per output is fixed).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int64_t M, n, nzH;
} hgeom;

static void crop_of(const int32_t *CA, int64_t nzH, int64_t M, int64_t c, int64_t *lo, int64_t *hi)
{
    if (CA) {
        *lo = (int64_t)CA[c] - 1;
        *hi = (int64_t)CA[c + nzH] - 1; /* inclusive */
    } else {
        *lo = 0;
        *hi = M - 1;
    }
}

#define DEFINE_FORWARD(NAME, HT, VT, OT, ACC)                                                       \
    void NAME(const HT *H, const int64_t *hd, const int32_t *CA, const VT *V, int64_t X, int64_t Y,   \
              int64_t nz, OT *O, int nthreads)                                                       \
    {                                                                                                 \
        const int64_t M = hd[0], n = hd[2], nzH = hd[4];                                              \
        if (nthreads > 0) omp_set_num_threads_wrap(nthreads);                                        \
        _Pragma("omp parallel for schedule(static)")                                                  \
        for (int64_t py = 0; py < Y; ++py) {                                                          \
            ACC *col = (ACC *)calloc((size_t)X, sizeof(ACC));                                         \
            for (int64_t a = 0; a < n; ++a)                                                           \
                for (int64_t b = 0; b < n; ++b)                                                       \
                    for (int64_t c = 0; c < nz; ++c) {                                                \
                        int64_t lo, hi;                                                               \
                        crop_of(CA, nzH, M, c, &lo, &hi);                                             \
                        const int64_t Mc = hi - lo + 1, off = Mc / 2;                                 \
                        const HT *K = H + M * M * (a + n * (b + n * c));                              \
                        /* q = (qx, qy) on sublattice (a, b); j = py - qy + off in [0, Mc) */         \
                        int64_t qy0 = py + off - (Mc - 1), qy1 = py + off;                            \
                        if (qy0 < 0) qy0 = 0;                                                         \
                        if (qy1 > Y - 1) qy1 = Y - 1;                                                 \
                        int64_t qs = qy0 + ((b - qy0) % n + n) % n;                                   \
                        for (int64_t qy = qs; qy <= qy1; qy += n) {                                   \
                            const int64_t j = py - qy + off;                                          \
                            const HT *Kc = K + M * (lo + j) + lo;                                     \
                            const VT *Vc = V + X * (qy + Y * c);                                      \
                            for (int64_t qx = a; qx < X; qx += n) {                                   \
                                const ACC v = (ACC)Vc[qx];                                            \
                                if (v == 0) continue;                                                 \
                                int64_t px0 = qx - off, px1 = qx - off + Mc - 1;                      \
                                int64_t i0 = 0;                                                       \
                                if (px0 < 0) { i0 = -px0; px0 = 0; }                                  \
                                if (px1 > X - 1) px1 = X - 1;                                         \
                                for (int64_t px = px0, i = i0; px <= px1; ++px, ++i)                  \
                                    col[px] += v * (ACC)Kc[i];                                        \
                            }                                                                         \
                        }                                                                             \
                    }                                                                                 \
            for (int64_t px = 0; px < X; ++px) O[px + X * py] = (OT)col[px];                          \
            free(col);                                                                                \
        }                                                                                             \
    }

#define DEFINE_BACKWARD(NAME, HT, IT, OT, ACC)                                                      \
    void NAME(const HT *H, const int64_t *hd, const int32_t *CA, const IT *Oimg, int64_t X,          \
              int64_t Y, int64_t nz, OT *B, int nthreads)                                            \
    {                                                                                                 \
        const int64_t M = hd[0], n = hd[2], nzH = hd[4];                                              \
        if (nthreads > 0) omp_set_num_threads_wrap(nthreads);                                        \
        _Pragma("omp parallel for schedule(static) collapse(2)")                                      \
        for (int64_t c = 0; c < nz; ++c)                                                              \
            for (int64_t qy = 0; qy < Y; ++qy) {                                                      \
                int64_t lo, hi;                                                                       \
                crop_of(CA, nzH, M, c, &lo, &hi);                                                     \
                const int64_t Mc = hi - lo + 1, offb = Mc - 1 - Mc / 2;                               \
                const int64_t b = qy % n;                                                             \
                for (int64_t qx = 0; qx < X; ++qx) {                                                  \
                    const int64_t a = qx % n;                                                         \
                    const HT *K = H + M * M * (a + n * (b + n * c));                                  \
                    ACC s = 0;                                                                        \
                    for (int64_t j = 0; j < Mc; ++j) {                                                \
                        const int64_t py = qy + j - offb;                                             \
                        if (py < 0 || py >= Y) continue;                                              \
                        const HT *Kc = K + M * (lo + j) + lo;                                         \
                        const IT *Oc = Oimg + X * py;                                                 \
                        int64_t i0 = 0, i1 = Mc - 1;                                                  \
                        if (qx - offb < 0) i0 = offb - qx;                                            \
                        if (qx - offb + Mc - 1 > X - 1) i1 = X - 1 - qx + offb;                       \
                        _Pragma("omp simd reduction(+ : s)")                                          \
                        for (int64_t i = i0; i <= i1; ++i)                                            \
                            s += (ACC)Oc[qx + i - offb] * (ACC)Kc[i];                                 \
                    }                                                                                 \
                    B[qx + X * (qy + Y * c)] = (OT)s;                                                 \
                }                                                                                     \
            }                                                                                         \
    }

static void omp_set_num_threads_wrap(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* FP64 oracle. */
DEFINE_FORWARD(orc_forward_f64, double, double, double, double)
/* FP64 oracle on an FP32 PSF (exact FP64 products of the FP32 kernel values):
 * avoids an FP64 copy of multi-GB PSFs. */
DEFINE_FORWARD(orc_forward_h32, float, double, double, double)
DEFINE_BACKWARD(orc_backward_f64, double, double, double, double)
/* Reference CPU semantics: single-precision forward (forwardProjectACC.m:5,25),
 * single-precision convolution stored as double (backwardProjectACC_new.m:6,22). */
DEFINE_FORWARD(orc_forward_f32, float, double, float, float)
DEFINE_BACKWARD(orc_backward_f32, float, float, double, float)

/* deconvRL.m:46-52,85-87 in FP64 with the FP64 projectors above.
 * X holds Xguess on entry and the result on exit; work needs X*Y + X*Y*nz doubles. */
void orc_isra_f64(const double *H, const int64_t *hd, const int32_t *CA, const double *Htf, double *Xg,
                  int64_t X, int64_t Y, int64_t nz, int iters, double *work, int nthreads)
{
    double *img = work, *D = work + X * Y;
    const int64_t nv = X * Y * nz;
    for (int it = 0; it < iters; ++it) {
        orc_forward_f64(H, hd, CA, Xg, X, Y, nz, img, nthreads);
        orc_backward_f64(H, hd, CA, img, X, Y, nz, D, nthreads);
        for (int64_t i = 0; i < nv; ++i) {
            double r = Htf[i] / D[i];
            if (isinf(r)) r = 0.0;
            double x = Xg[i] * r;
            if (isnan(x)) x = 0.0;
            Xg[i] = x;
        }
    }
}

/* Same loop with the reference CPU path's precisions (reconstruct_S.m:118-121
 * with forwardProjectACC / backwardProjectACC_new): the volume and Htf are
 * double, the forward converts the masked plane to single and accumulates in
 * single, the backward convolves in single and stores double.  This is the
 * CPU baseline.  img needs X*Y floats, D needs X*Y*nz doubles. */
void orc_isra_f32(const float *H, const int64_t *hd, const int32_t *CA, const double *Htf, double *Xg,
                  int64_t X, int64_t Y, int64_t nz, int iters, float *img, double *D, int nthreads)
{
    const int64_t nv = X * Y * nz;
    for (int it = 0; it < iters; ++it) {
        orc_forward_f32(H, hd, CA, Xg, X, Y, nz, img, nthreads);
        orc_backward_f32(H, hd, CA, img, X, Y, nz, D, nthreads);
        for (int64_t i = 0; i < nv; ++i) {
            double r = Htf[i] / D[i];
            if (isinf(r)) r = 0.0;
            double x = Xg[i] * r;
            if (isnan(x)) x = 0.0;
            Xg[i] = x;
        }
    }
}

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
