// Background and standard-deviation image of a movie (the producer of the RL
// input in sid_main.m:88-103): optional division by the global maximum
// (sid_main.m:88-89), rank_1_factorization (src/rank_1_factorization.m:31-39)
// and compute_std_image (src/compute_std_image.m:22-26).
//
// The movie Y is column-major [P pixels, T frames] in HBM (one frame per
// column), read-only.  Every step is one streaming pass over Y:
//   column pass  t_j = sum_i s_i Y_ij            (block per frame)
//   row pass     s_i = sum_j Y_ij t_j            (thread per pixel; in the last
//                A_i = sum_j Y_ij^2, Sum_i = sum_j Y_ij      iteration also these)
// with FP64 accumulation and fixed-order reductions (deterministic).  Norms of
// s and t are reduced on the host in FP64.  The normalising maximum comes out of
// the first column pass (the first t is normalised, so its scale is applied on
// the host afterwards), or out of a pass of its own without background
// iterations: 2 * bg_iter passes over Y in all (4.9 -> 3.3 ms at 1010025 px x
// 1000 frames with the maximum fused).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "sid_internal.h"

namespace sid {

constexpr int kMvThreads = 256;

__global__ __launch_bounds__(kMvThreads) void k_mv_max(const float *__restrict__ Y, int64_t n, float *__restrict__ part)
{
    __shared__ float red[kMvThreads];
    float m = -INFINITY;
    // four independent loads per trip keep enough bytes in flight (max is order-independent)
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; e + 3 * stride < n; e += 4 * stride) {
        const float a = Y[e], b = Y[e + stride], c = Y[e + 2 * stride], d = Y[e + 3 * stride];
        m = fmaxf(m, fmaxf(fmaxf(a, b), fmaxf(c, d)));
    }
    for (; e < n; e += stride) m = fmaxf(m, Y[e]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = kMvThreads / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Column pass over frames j0..j0+3 and pixel chunk c:
// part[c][j] = scale * sum_{i in chunk c} s_i * Y[i, j], s_i = sd[i] (or 1 when sd is null).
// Four frames per block so s is read once per four columns; 8 or 16 frames per
// block (1-4 pixels per trip, shuffle reductions) measured 8-100 % slower for
// the pass although they read s less often.  Chunks keep the grid many blocks
// deep; the host adds the chunks in order.
constexpr int kColGroup = 4;
// With mpart the pass also writes the block's maximum of Y (the global maximum
// of the normalisation then costs no pass of its own).
__global__ __launch_bounds__(kMvThreads) void k_mv_col(const float *__restrict__ Y, int64_t P, int64_t T,
                                                       const double *__restrict__ sd, double scale, int64_t pc,
                                                       double *__restrict__ part, float *__restrict__ mpart)
{
    __shared__ double red[kColGroup][kMvThreads];
    __shared__ float redm[kMvThreads];
    float mx = -INFINITY;
    const int64_t j0 = (int64_t)blockIdx.x * kColGroup;
    const int64_t i0 = (int64_t)blockIdx.y * pc, i1 = std::min(P, i0 + pc);
    const int nc = (int)std::min<int64_t>(kColGroup, T - j0);
    double acc[kColGroup];
#pragma unroll
    for (int c = 0; c < kColGroup; ++c) acc[c] = 0.0;
    // kColU pixels per trip: their 4 x kColU loads are in flight together; each
    // thread still adds its pixels in ascending order
    constexpr int kColU = 4;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kColU * kMvThreads) {
        double sv[kColU];
        float y[kColU][kColGroup];
#pragma unroll
        for (int u = 0; u < kColU; ++u) {
            const int64_t ii = i + (int64_t)u * kMvThreads;
            const bool in = ii < i1;
            sv[u] = in ? (sd ? sd[ii] : 1.0) : 0.0;
#pragma unroll
            for (int c = 0; c < kColGroup; ++c) {
                y[u][c] = (in && c < nc) ? Y[(j0 + c) * P + ii] : 0.0f;
                if (in && c < nc) mx = fmaxf(mx, y[u][c]);
            }
        }
#pragma unroll
        for (int u = 0; u < kColU; ++u) {
            if (i + (int64_t)u * kMvThreads >= i1) break;
#pragma unroll
            for (int c = 0; c < kColGroup; ++c) acc[c] += (double)y[u][c] * sv[u];
        }
    }
#pragma unroll
    for (int c = 0; c < kColGroup; ++c) red[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int o = kMvThreads / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
#pragma unroll
            for (int c = 0; c < kColGroup; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + o];
        __syncthreads();
    }
    if ((int)threadIdx.x < nc) part[(int64_t)blockIdx.y * T + j0 + threadIdx.x] = red[threadIdx.x][0] * scale;
    if (mpart) {   // uniform
        redm[threadIdx.x] = mx;
        __syncthreads();
        for (int o = kMvThreads / 2; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) redm[threadIdx.x] = fmaxf(redm[threadIdx.x], redm[threadIdx.x + o]);
            __syncthreads();
        }
        if (threadIdx.x == 0) mpart[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = redm[0];
    }
}

// Row pass: sd[i] = scale * sum_j Y[i, j] * t[j] (when t), and, when stats,
// A[i] = scale^2 * sum_j Y[i, j]^2, Sm[i] = scale * sum_j Y[i, j].
__global__ __launch_bounds__(kMvThreads) void k_mv_row(const float *__restrict__ Y, int64_t P, int64_t T,
                                                       const double *__restrict__ t, double scale,
                                                       double *__restrict__ sd, double *__restrict__ A,
                                                       double *__restrict__ Sm)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    double s = 0.0, a = 0.0, m = 0.0;
    const float *p = Y + i;
    int64_t j = 0;
    constexpr int U = 8;
    for (; j + U <= T; j += U) {
        float y[U];
#pragma unroll
        for (int u = 0; u < U; ++u) y[u] = p[(j + u) * P];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const double v = y[u];
            if (t) s += v * t[j + u];
            if (A) {
                a += v * v;
                m += v;
            }
        }
    }
    for (; j < T; ++j) {
        const double v = p[j * P];
        if (t) s += v * t[j];
        if (A) {
            a += v * v;
            m += v;
        }
    }
    if (sd) sd[i] = s * scale;
    if (A) {
        A[i] = a * scale * scale;
        Sm[i] = m * scale;
    }
}

__global__ __launch_bounds__(kMvThreads) void k_mv_sumsq(const double *__restrict__ x, int64_t n,
                                                         double *__restrict__ part)
{
    __shared__ double red[kMvThreads];
    double a = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        a += x[e] * x[e];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int o = kMvThreads / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// compute_std_image.m:26 with y1 = s (sd * s_scale), C = mean(t):
//   std = sqrt((A - y1^2) / T - B^2 + 2 B y1 C - y1^2 C^2),  B = Sm / T.
// A negative radicand (rounding, constant pixels) gives 0.
__global__ void k_mv_std(const double *__restrict__ A, const double *__restrict__ Sm, const double *__restrict__ sd,
                         double s_scale, int64_t P, int64_t T, double C, float *__restrict__ std_img,
                         float *__restrict__ s_out)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < P; i += (int64_t)gridDim.x * blockDim.x) {
        const double y1 = sd ? sd[i] * s_scale : 0.0;
        const double B = Sm[i] / (double)T;
        const double v = (A[i] - y1 * y1) / (double)T - B * B + 2.0 * B * y1 * C - y1 * y1 * C * C;
        std_img[i] = (float)sqrt(v > 0.0 ? v : 0.0);
        if (s_out) s_out[i] = (float)y1;
    }
}

static int grid_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096)); }

// Column-pass geometry: groups of kColGroup frames by pixel chunks, about
// 16 K blocks in all (many rounds of the ~2 K resident ones, so no half-empty
// last round).
static void col_geometry(int64_t P, int64_t T, int64_t *ngroups, int64_t *nchunk, int64_t *pc)
{
    *ngroups = (T + kColGroup - 1) / kColGroup;
    *nchunk = std::max<int64_t>(1, std::min<int64_t>((16384 + *ngroups - 1) / *ngroups, (P + 1023) / 1024));
    *pc = (P + *nchunk - 1) / *nchunk;
}

// Device work space of one call (partial sums, the FP64 s and t vectors and
// per-pixel sums), carved out of one buffer at 256-byte boundaries.
struct MvWork {
    float *fpart = nullptr, *mpart = nullptr;
    double *dpart = nullptr, *A = nullptr, *Sm = nullptr, *sd = nullptr, *td = nullptr;
};

// Bytes of work space for (P, T); with base != nullptr also sets the pointers.
size_t mv_work_bytes(int64_t P, int64_t T, MvWork *w, void *base)
{
    int64_t ngroups, nchunk, pc;
    col_geometry(P, T, &ngroups, &nchunk, &pc);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) / 256 * 256;
        return base ? (void *)((char *)base + o) : nullptr;
    };
    MvWork t;
    t.fpart = (float *)take(sizeof(float) * 4096);
    t.dpart = (double *)take(sizeof(double) * 4096);
    t.A = (double *)take(sizeof(double) * P);
    t.Sm = (double *)take(sizeof(double) * P);
    t.sd = (double *)take(sizeof(double) * P);
    t.td = (double *)take(sizeof(double) * std::max<int64_t>(T, nchunk * T));
    t.mpart = (float *)take(sizeof(float) * ngroups * nchunk);
    if (w) *w = t;
    return off;
}

static int host_sumsq(const double *x, int64_t n, double *part, hipStream_t st, double *out)
{
    const int nb = grid_for(n);
    hipLaunchKernelGGL(k_mv_sumsq, dim3(nb), dim3(kMvThreads), 0, st, x, n, part);
    SID_TRY_HIP(hipGetLastError());
    std::vector<double> h(nb);
    SID_TRY_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
    SID_TRY_HIP(hipStreamSynchronize(st));
    double s = 0.0;
    for (double v : h) s += v;
    *out = s;
    return SID_OK;
}

// work: mv_work_bytes(P, T) bytes of device memory (every value in it is
// written before it is read within the call).
int std_image(const float *Y, int64_t P, int64_t T, int bg_iter, int flags, float *s_out, float *t_out,
              float *std_img, double *ymax_out, hipStream_t st, void *work)
{
    MvWork w;
    mv_work_bytes(P, T, &w, work);
    // 1. global maximum (sid_main.m:88-89), when normalising; with background
    // iterations it comes out of the first column pass (below)
    double scale = 1.0, ymax = 0.0;
    const bool fuse_max = (flags & 1) && bg_iter > 0;
    if ((flags & 1) && !fuse_max) {
        const int nb = grid_for(P * T);
        hipLaunchKernelGGL(k_mv_max, dim3(nb), dim3(kMvThreads), 0, st, Y, P * T, w.fpart);
        SID_TRY_HIP(hipGetLastError());
        std::vector<float> h(nb);
        SID_TRY_HIP(hipMemcpyAsync(h.data(), w.fpart, sizeof(float) * nb, hipMemcpyDeviceToHost, st));
        SID_TRY_HIP(hipStreamSynchronize(st));
        float m = -INFINITY;
        for (float v : h) m = std::max(m, v);
        ymax = m;
        if (!(ymax > 0.0) || !std::isfinite(ymax)) {
            set_error("movie maximum %g: cannot normalise", ymax);
            return SID_ERR_ARG;
        }
        scale = 1.0 / ymax;
    }
    if (ymax_out) *ymax_out = ymax;
    double s_scale = 0.0, C = 0.0;
    std::vector<double> th(T, 0.0);
    if (bg_iter > 0) {
        int64_t ngroups, nchunk, pc;
        col_geometry(P, T, &ngroups, &nchunk, &pc);
        std::vector<double> ph(nchunk * T);
        // s = ones(P, 1) / sqrt(P) (rank_1_factorization.m:31)
        double s_mul = 1.0 / std::sqrt((double)P);
        bool s_const = true;
        for (int it = 1; it <= bg_iter; ++it) {
            // t = s' * Y; t = t / norm(t)   (:33-34)
            // the first pass of a normalised movie finds the maximum and leaves the
            // chunk sums unscaled; they are scaled here by the same factor the
            // kernel would apply (identical products)
            const bool find_max = fuse_max && it == 1;
            hipLaunchKernelGGL(k_mv_col, dim3((unsigned)ngroups, (unsigned)nchunk), dim3(kMvThreads), 0, st, Y, P, T,
                               s_const ? (const double *)nullptr : w.sd, find_max ? 1.0 : s_mul * scale, pc, w.td,
                               find_max ? w.mpart : (float *)nullptr);
            SID_TRY_HIP(hipGetLastError());
            SID_TRY_HIP(hipMemcpyAsync(ph.data(), w.td, sizeof(double) * nchunk * T, hipMemcpyDeviceToHost, st));
            if (find_max) {
                std::vector<float> h(ngroups * nchunk);
                SID_TRY_HIP(hipMemcpyAsync(h.data(), w.mpart, sizeof(float) * h.size(), hipMemcpyDeviceToHost, st));
                SID_TRY_HIP(hipStreamSynchronize(st));
                float m = -INFINITY;
                for (float v : h) m = std::max(m, v);
                ymax = m;
                if (!(ymax > 0.0) || !std::isfinite(ymax)) {
                    set_error("movie maximum %g: cannot normalise", ymax);
                    return SID_ERR_ARG;
                }
                scale = 1.0 / ymax;
                if (ymax_out) *ymax_out = ymax;
                const double f = s_mul * scale;
                for (double &v : ph) v *= f;
            }
            SID_TRY_HIP(hipStreamSynchronize(st));
            double nt = 0.0;
            for (int64_t j = 0; j < T; ++j) {
                double v = 0.0;
                for (int64_t c = 0; c < nchunk; ++c) v += ph[c * T + j];
                th[j] = v;
                nt += v * v;
            }
            nt = std::sqrt(nt);
            for (double &v : th) v /= nt;   // 0 / 0 = NaN, as MATLAB
            SID_TRY_HIP(hipMemcpyAsync(w.td, th.data(), sizeof(double) * T, hipMemcpyHostToDevice, st));
            // s = Y * t' (:35), with the std-image sums in the last iteration
            const bool last = it == bg_iter;
            hipLaunchKernelGGL(k_mv_row, dim3((unsigned)((P + kMvThreads - 1) / kMvThreads)), dim3(kMvThreads), 0, st,
                               Y, P, T, (const double *)w.td, scale, w.sd, last ? w.A : (double *)nullptr,
                               last ? w.Sm : (double *)nullptr);
            SID_TRY_HIP(hipGetLastError());
            s_const = false;
            if (!last) {   // s = s / norm(s) (:36-38), folded into the next column pass
                double ns = 0.0;
                SID_TRY(host_sumsq(w.sd, P, w.dpart, st, &ns));
                s_mul = 1.0 / std::sqrt(ns);
            }
        }
        s_scale = 1.0;
        double sum_t = 0.0;
        for (double v : th) sum_t += v;
        C = sum_t / (double)T;
    } else {
        hipLaunchKernelGGL(k_mv_row, dim3((unsigned)((P + kMvThreads - 1) / kMvThreads)), dim3(kMvThreads), 0, st, Y,
                           P, T, (const double *)nullptr, scale, (double *)nullptr, w.A, w.Sm);
        SID_TRY_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_mv_std, dim3(grid_for(P)), dim3(256), 0, st, w.A, w.Sm,
                       bg_iter > 0 ? (const double *)w.sd : (const double *)nullptr, s_scale, P, T, C, std_img,
                       s_out);
    SID_TRY_HIP(hipGetLastError());
    if (t_out) {
        std::vector<float> tf(T);
        for (int64_t j = 0; j < T; ++j) tf[j] = (float)th[j];
        SID_TRY_HIP(hipMemcpyAsync(t_out, tf.data(), sizeof(float) * T, hipMemcpyHostToDevice, st));
    }
    SID_TRY_HIP(hipStreamSynchronize(st));
    return SID_OK;
}

// ---------------------------------------------------------------------------
// Work space without allocation (sid_std_image_device_noalloc): a file-scope
// device array, one copy per device, looked up once per device.  Calls that
// overlap on one device share it and race.
// ---------------------------------------------------------------------------
constexpr size_t kMvScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) unsigned char scratch[kMvScratchBytes];
constexpr int kMaxDevices = 64;
static void *g_scratch[kMaxDevices] = {};
static std::mutex g_scratch_mu;

size_t mv_scratch_bytes() { return kMvScratchBytes; }

int std_image_noalloc(const float *Y, int64_t P, int64_t T, int bg_iter, int flags, float *s_out, float *t_out,
                      float *std_img, double *ymax_out, hipStream_t st)
{
    const size_t need = mv_work_bytes(P, T, nullptr, nullptr);
    if (need > kMvScratchBytes) {
        set_error("movie of %lld px x %lld frames needs %zu bytes of work space; the allocation-free entry holds %zu",
                  (long long)P, (long long)T, need, kMvScratchBytes);
        return SID_ERR_DIMS;
    }
    int dev = 0;
    SID_TRY_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device %d beyond the %d devices the allocation-free entry supports", dev, kMaxDevices);
        return SID_ERR_NODEV;
    }
    void *ptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch[dev]) {
            void *q = nullptr;
            SID_TRY_HIP(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
            g_scratch[dev] = q;
        }
        ptr = g_scratch[dev];
    }
    return std_image(Y, P, T, bg_iter, flags, s_out, t_out, std_img, ymax_out, st, ptr);
}

}  // namespace sid
