// LS_deconv on the device: external/LFrecon_util/LS_deconv.m:1-97.
// Vectors are column-major device volumes {X, Y, Nz}; the projector passes
// are the sid_forward/sid_backward device passes; dot products are FP64
// block partials summed on the host in a fixed order (deterministic).
//
//   Q(x) = B(F(x)) + tv(x) + lamb_L2 x                                  (:67-73)
//   X = df;  g = Q(df) - df + lamb_L1                                  (:74-75)
//   repeat max_iter times                                              (:78-96)
//     d = ((X > 0) | (g < 0)) .* g
//     alpha = d'd / ||F(d)||^2   or  d'd / d'Q(d) (all three products nonzero, :80-84)
//     alpha = 0 when NaN or Inf
//     Xn = max(X - alpha d, 0);  g = g + Q(Xn - X);  X = Xn
//
// tv(x) is convn(x, filter, 'same') with the 3x3x3 filter of :54-58:
//   tv(x)[i,j,k] = l1 (x[i-1,j,k] + x[i+1,j,k]) + l2 (x[i,j-1,k] + x[i,j+1,k])
//                + l3 (x[i,j,k-1] + x[i,j,k+1]) - (l1 + l2 + l3) x[i,j,k]
// with zeros outside the volume (the filter is symmetric, so convolution and
// correlation agree).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "sid_internal.h"

namespace sid {

struct LsTerms {
    int64_t nx, ny, nz;
    float l1, l2, l3;  // TV_L2 multipliers per dimension (0 when the term is off)
    float lamb_l2;     // L2 multiplier (0 when the term is off)
    int tv;            // tv term present
};

__device__ __forceinline__ float extra_terms(const float *__restrict__ x, int64_t e, const LsTerms &t)
{
    const float xv = x[e];
    float r = t.lamb_l2 * xv;
    if (t.tv) {
        const int64_t i = e % t.nx, jk = e / t.nx, j = jk % t.ny, k = jk / t.ny;
        const int64_t sy = t.nx, sz = t.nx * t.ny;
        float s = -(t.l1 + t.l2 + t.l3) * xv;
        if (i > 0) s += t.l1 * x[e - 1];
        if (i + 1 < t.nx) s += t.l1 * x[e + 1];
        if (j > 0) s += t.l2 * x[e - sy];
        if (j + 1 < t.ny) s += t.l2 * x[e + sy];
        if (k > 0) s += t.l3 * x[e - sz];
        if (k + 1 < t.nz) s += t.l3 * x[e + sz];
        r += s;
    }
    return r;
}

// q = B(F(src)) (bq) plus the penalty terms of src, combined into out:
//   mode 0: out = q - src + c      (initial gradient, :75)
//   mode 1: out = out + q          (gradient update, :94)
//   mode 2: out = q                (Q(d) for the alternate step, :81)
__global__ void k_ls_qcomb(const float *bq, const float *__restrict__ src, float *out,
                           int mode, float c, LsTerms t)
{
    const int64_t n = t.nx * t.ny * t.nz;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const float q = bq[e] + extra_terms(src, e, t);
        if (mode == 0)
            out[e] = q - src[e] + c;
        else if (mode == 1)
            out[e] = out[e] + q;
        else
            out[e] = q;
    }
}

// d = passive .* g with passive = (X > 0) | (g < 0)   (:79-80)
__global__ void k_ls_dir(const float *__restrict__ X, const float *__restrict__ g, int64_t n, float *__restrict__ d)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const float x = X[e], gv = g[e];
        d[e] = (x > 0.0f || gv < 0.0f) ? gv : 0.0f;
    }
}

// Xn = max(X - alpha d, 0) written over d; t = Xn - X   (:89-91, 94)
__global__ void k_ls_step(const float *__restrict__ X, float *__restrict__ d, const float *__restrict__ alpha_p,
                          int64_t n, float *__restrict__ t)
{
    const float alpha = *alpha_p;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        float xn = X[e] - alpha * d[e];
        if (xn < 0.0f) xn = 0.0f;
        d[e] = xn;
        t[e] = xn - X[e];
    }
}

constexpr int kDotBlock = 256;
constexpr int kDotMaxBlocks = 2048;

__global__ void __launch_bounds__(kDotBlock) k_dot_partial(const float *__restrict__ a, const float *__restrict__ b,
                                                           int64_t n, double *__restrict__ partial)
{
    __shared__ double red[kDotBlock];
    double s = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        s += (double)a[e] * (double)b[e];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = kDotBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

static int nblk(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, kDotMaxBlocks)); }

// alpha = num / den (:84), 0 when NaN or Inf (:85-87), from the block partials
// of the two dot products summed in block order (one thread, the same order a
// host sum would take), so the step size never leaves the device.
__global__ void __launch_bounds__(kDotBlock) k_ls_alpha(const double *__restrict__ pn, int nn,
                                                        const double *__restrict__ pd, int nd, float *__restrict__ alpha)
{
    __shared__ double sn[kDotMaxBlocks], sd[kDotMaxBlocks];
    for (int i = threadIdx.x; i < nn; i += blockDim.x) sn[i] = pn[i];
    for (int i = threadIdx.x; i < nd; i += blockDim.x) sd[i] = pd[i];
    __syncthreads();
    if (threadIdx.x != 0) return;
    double num = 0.0, den = 0.0;
    for (int i = 0; i < nn; ++i) num += sn[i];
    for (int i = 0; i < nd; ++i) den += sd[i];
    double a = num / den;
    if (std::isnan(a) || std::isinf(a)) a = 0.0;
    *alpha = (float)a;
}

struct Bufs {
    std::vector<void *> ptrs;
    ~Bufs()
    {
        for (void *q : ptrs)
            if (q) (void)hipFree(q);
    }
    template <typename T> int alloc(T **p, int64_t n)
    {
        void *q = nullptr;
        const hipError_t e = hipMalloc(&q, sizeof(T) * std::max<int64_t>(n, 1));
        if (e != hipSuccess) {
            set_error("hipMalloc failed: %s", hipGetErrorString(e));
            return SID_ERR_OOM;
        }
        ptrs.push_back(q);
        *p = (T *)q;
        return SID_OK;
    }
};

// FP64 block partials of a'b (nblk(n) of them, into partial)
static int dot_partials(const float *a, const float *b, int64_t n, double *partial, hipStream_t st)
{
    hipLaunchKernelGGL(k_dot_partial, dim3(nblk(n)), dim3(kDotBlock), 0, st, a, b, n, partial);
    SID_TRY_HIP(hipGetLastError());
    return SID_OK;
}

int ls_deconv(sid_psf *p, const float *df, float *Xout, const int64_t vdims[3], const sid_ls_opts &o, hipStream_t st)
{
    const int64_t nv = vdims[0] * vdims[1] * vdims[2], ni = vdims[0] * vdims[1];
    const int64_t odims[2] = {vdims[0], vdims[1]};
    LsTerms t{};
    t.nx = vdims[0];
    t.ny = vdims[1];
    t.nz = vdims[2];
    // `if max(opts.lamb_TV_L2)` (:50) and `if opts.lamb_L2` (:67)
    const double tvmax = std::max(o.lamb_TV_L2[0], std::max(o.lamb_TV_L2[1], o.lamb_TV_L2[2]));
    t.tv = tvmax != 0.0;
    if (t.tv) {
        t.l1 = (float)o.lamb_TV_L2[0];
        t.l2 = (float)o.lamb_TV_L2[1];
        t.l3 = (float)o.lamb_TV_L2[2];
    }
    t.lamb_l2 = (float)o.lamb_L2;
    // `if opts.lamb_TV_L1*opts.lamb_TV_L2*opts.lamb_L2` (:80): true when every entry is nonzero
    bool alt = true;
    for (int i = 0; i < 3; ++i) alt = alt && (o.lamb_TV_L1 * o.lamb_TV_L2[i] * o.lamb_L2 != 0.0);

    Bufs b;
    float *X, *g, *d, *tt, *q, *img, *alpha;
    double *pnum, *pden;
    SID_TRY(b.alloc(&X, nv));
    SID_TRY(b.alloc(&g, nv));
    SID_TRY(b.alloc(&d, nv));
    SID_TRY(b.alloc(&tt, nv));
    SID_TRY(b.alloc(&q, nv));
    SID_TRY(b.alloc(&img, ni));
    SID_TRY(b.alloc(&pnum, kDotMaxBlocks));
    SID_TRY(b.alloc(&pden, kDotMaxBlocks));
    SID_TRY(b.alloc(&alpha, 1));
    const int nbv = nblk(nv);
    // every pass below stays queued on st: the projections skip their host waits
    // and the step size is computed on the device
    SID_TRY_HIP(hipMemcpyAsync(X, df, sizeof(float) * nv, hipMemcpyDeviceToDevice, st));
    // g = Q(df) - df + lamb_L1
    SID_TRY(forward_dev(p, X, vdims, img, st, false));
    SID_TRY(backward_dev(p, img, odims, q, st, false));
    hipLaunchKernelGGL(k_ls_qcomb, dim3(nbv), dim3(256), 0, st, q, X, g, 0, (float)o.lamb_L1, t);
    SID_TRY_HIP(hipGetLastError());
    for (int it = 0; it < o.max_iter; ++it) {
        hipLaunchKernelGGL(k_ls_dir, dim3(nbv), dim3(256), 0, st, X, g, nv, d);
        SID_TRY_HIP(hipGetLastError());
        SID_TRY(forward_dev(p, d, vdims, img, st, false));
        int nden;
        if (alt) {
            SID_TRY(backward_dev(p, img, odims, q, st, false));
            hipLaunchKernelGGL(k_ls_qcomb, dim3(nbv), dim3(256), 0, st, q, d, q, 2, 0.0f, t);
            SID_TRY_HIP(hipGetLastError());
            SID_TRY(dot_partials(d, q, nv, pden, st));
            nden = nbv;
        } else {
            SID_TRY(dot_partials(img, img, ni, pden, st));
            nden = nblk(ni);
        }
        SID_TRY(dot_partials(d, d, nv, pnum, st));
        hipLaunchKernelGGL(k_ls_alpha, dim3(1), dim3(kDotBlock), 0, st, pnum, nbv, pden, nden, alpha);
        SID_TRY_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_ls_step, dim3(nbv), dim3(256), 0, st, X, d, alpha, nv, tt);
        SID_TRY_HIP(hipGetLastError());
        // g = g + Q(Xn - X)
        SID_TRY(forward_dev(p, tt, vdims, img, st, false));
        SID_TRY(backward_dev(p, img, odims, q, st, false));
        hipLaunchKernelGGL(k_ls_qcomb, dim3(nbv), dim3(256), 0, st, q, tt, g, 1, 0.0f, t);
        SID_TRY_HIP(hipGetLastError());
        std::swap(X, d);  // X = Xn
    }
    SID_TRY_HIP(hipMemcpyAsync(Xout, X, sizeof(float) * nv, hipMemcpyDeviceToDevice, st));
    SID_TRY_HIP(hipStreamSynchronize(st));
    double f_ms, b_ms;
    fft_timing_collect(p, &f_ms, &b_ms);   // release the FFT path's timing events
    return SID_OK;
}

// ISRA update of deconvRL.m:85-87 on column-major buffers:
// R = Htf ./ D; R(isinf(R)) = 0; X = X .* R; X(isnan(X)) = 0.
__global__ void k_rl_update(float *__restrict__ X, const float *__restrict__ htf, const float *__restrict__ D,
                            int64_t n)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        float r = htf[e] / D[e];
        if (isinf(r)) r = 0.0f;
        float v = X[e] * r;
        if (isnan(v)) v = 0.0f;
        X[e] = v;
    }
}

// deconvRL basic (lambda = 0) with the handle's wrapped projectors
// (forwardFUN = F(convn(X, k)), backwardFUN = convn(B(y), k)): deconvRL.m:46-52,85-92.
int rl_wrapped(sid_psf *p, const float *Htf, float *X, const int64_t vdims[3], int max_iter, float *hist,
               hipStream_t st)
{
    const int64_t nv = vdims[0] * vdims[1] * vdims[2], ni = vdims[0] * vdims[1];
    const int64_t odims[2] = {vdims[0], vdims[1]};
    Bufs b;
    float *img, *D;
    SID_TRY(b.alloc(&img, ni));
    SID_TRY(b.alloc(&D, nv));
    for (int it = 0; it < max_iter; ++it) {   // queued on st without host waits
        SID_TRY(forward_dev(p, X, vdims, img, st, false));
        SID_TRY(backward_dev(p, img, odims, D, st, false));
        hipLaunchKernelGGL(k_rl_update, dim3(nblk(nv)), dim3(256), 0, st, X, Htf, D, nv);
        SID_TRY_HIP(hipGetLastError());
        if (hist) SID_TRY_HIP(hipMemcpyAsync(hist + (int64_t)it * nv, X, sizeof(float) * nv, hipMemcpyDeviceToDevice, st));
    }
    SID_TRY_HIP(hipStreamSynchronize(st));
    double f_ms, b_ms;
    fft_timing_collect(p, &f_ms, &b_ms);   // release the FFT path's timing events
    return SID_OK;
}

}  // namespace sid
