// Poisson log-likelihood of the light-field forward model and its exact
// gradients (not in the reference; used for PSF refinement and as the
// adjoint check of the projector pair).
//
//   NLL(V, H) = sum_f sum_p F_f[p] - y_f[p] log F_f[p],   F_f = forward(H, V_f)
//   dNLL/dV_f = backward(H, r_f),                          r_f = 1 - y_f / F_f
//   dNLL/dH   = sum_f  A_f^T R_f   per forward GEMM segment (plane c, tap row dx)
//
// The last line is the weight gradient of the forward implicit GEMM
// C = A W (sid_direct.hip): for segment (c, dx) the Toeplitz A rows of the
// volume give dW[k][uv] = sum_{f,P} A_f[P][k] r_f[P][uv], with the K index
// k = (t_hi - dy) * S + ab of the forward weights; a gather maps dW back to
// the MATLAB layout of H.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "sid_internal.h"

namespace sid {

// r = 1 - y / F on the padded image (in place) and per-block FP64 partial sums
// of F - y log F.  y == 0 contributes F with r = 1 (the limit of y log F).
// Phantom pixels (beyond X, Y) and padding stay 0.
__global__ void k_nll_residual(float *__restrict__ img, const float *__restrict__ y_col, int X, int Y, int n,
                               int64_t YP, int T, int S, double *__restrict__ partial)
{
    __shared__ double red[256];
    const int64_t total = (int64_t)X * Y;
    double acc = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(e % X), yy = (int)(e / X);
        const int Px = x / n, Py = yy / n;
        const int uv = (x - Px * n) + n * (yy - Py * n);
        const int64_t o = ((int64_t)(Px + T) * YP + (Py + T)) * S + uv;
        const double F = img[o];
        const double yv = y_col[e];
        double term, r;
        if (yv == 0.0) {
            term = F;
            r = 1.0;
        } else {
            term = F - yv * log(F);
            r = 1.0 - yv / F;
        }
        acc += term;
        img[o] = (float)r;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// dW for one (segment, row tile of K): C[k][uv] = sum_{f, P} A_f[P][k] R_f[P][uv].
// 256 threads = 4 waves, tile 128 K-rows x NP columns, the sum over (f, P) in
// chunks of 32 lenslets staged in LDS.  v_mfma_f32_16x16x4_f32 with the
// lenslet index as the MFMA K dimension: lane (l16, g) feeds A^T[row l16][P =
// 4s + g] and R[P = 4s + g][col l16] at step s; with LDS row strides of 16
// (mod 64) floats the four lane groups of a read hit disjoint banks.  The next
// chunk's tiles are loaded into registers while the current one is multiplied
// (one LDS buffer, two barriers per chunk).
struct DwItem {
    int32_t seg, ktile;
};

template <int NT>
__global__ __launch_bounds__(256, 2) void k_dw_gemm(const Seg *__restrict__ segs, const DwItem *__restrict__ items,
                                                    const float *__restrict__ vols, int64_t vol_stride,
                                                    const float *__restrict__ res, int64_t img_stride, int nb,
                                                    int nrows, int Ly, int64_t YP, int T, int S,
                                                    float *__restrict__ dW, const int64_t *__restrict__ dw_off)
{
    constexpr int NP = NT * 16;
    constexpr int MT = 128, PC = 32;
    constexpr int LDA = MT + 16, LDR = NP + (80 - NP % 64) % 64;
    constexpr int A4 = MT / 4, R4 = NP / 4;               // float4 per lenslet row of each tile
    constexpr int NA = A4 / 8, NR = (R4 + 7) / 8;         // 8 threads per lenslet row
    static_assert(PC * 8 == 256 && A4 % 8 == 0, "one lenslet row per 8 threads");
    static_assert(LDA % 64 == 16 && LDR % 64 == 16, "row strides of 16 (mod 64) floats");
    __shared__ float As[PC * LDA];
    __shared__ float Rs[PC * LDR];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, g = lane >> 4;
    const int pl = tid >> 3, q = tid & 7;                 // this thread's lenslet row in a chunk and column phase
    const DwItem it = items[blockIdx.x];
    const Seg sg = segs[it.seg];
    const int k0 = it.ktile * MT;
    const int cpf = (nrows + PC - 1) / PC;
    const int nch = nb * cpf;
    f4 acc[2][NT];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) acc[rt][ct] = f4{0.f, 0.f, 0.f, 0.f};

    float4 ra[NA], rr[NR];
    auto fetch = [&](int ci) {
        const int f = ci / cpf, r = (ci - f * cpf) * PC + pl;
#pragma unroll
        for (int j = 0; j < NA; ++j) ra[j] = float4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int j = 0; j < NR; ++j) rr[j] = float4{0.f, 0.f, 0.f, 0.f};
        if (r < nrows) {
            const int Px = r / Ly, Py = r - Px * Ly;
            const float *V = vols + (int64_t)f * vol_stride + sg.a_off + k0 + ((int64_t)Px * YP + Py) * S;
            const float *R = res + (int64_t)f * img_stride + ((int64_t)(Px + T) * YP + (Py + T)) * S;
#pragma unroll
            for (int j = 0; j < NA; ++j) ra[j] = *(const float4 *)(V + 4 * (q + 8 * j));
#pragma unroll
            for (int j = 0; j < NR; ++j) {
                const int c4 = q + 8 * j;
                if (c4 < R4 && 4 * c4 < S) rr[j] = *(const float4 *)(R + 4 * c4);
            }
        }
    };
    fetch(0);
    for (int ci = 0; ci < nch; ++ci) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NA; ++j) *(float4 *)(As + pl * LDA + 4 * (q + 8 * j)) = ra[j];
#pragma unroll
        for (int j = 0; j < NR; ++j)
            if (q + 8 * j < R4) *(float4 *)(Rs + pl * LDR + 4 * (q + 8 * j)) = rr[j];
        __syncthreads();
        if (ci + 1 < nch) fetch(ci + 1);
#pragma unroll
        for (int s = 0; s < PC / 4; ++s) {
            const int P = 4 * s + g;
            float a[2];
#pragma unroll
            for (int rt = 0; rt < 2; ++rt) a[rt] = As[P * LDA + wave * 32 + rt * 16 + l16];
#pragma unroll
            for (int ct = 0; ct < NT; ++ct) {
                const float b = Rs[P * LDR + ct * 16 + l16];
#pragma unroll
                for (int rt = 0; rt < 2; ++rt)
                    acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rt], b, acc[rt][ct], 0, 0, 0);
            }
        }
    }
    float *out = dW + dw_off[it.seg];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt) {
        const int rbase = k0 + wave * 32 + rt * 16 + 4 * g;
#pragma unroll
        for (int ct = 0; ct < NT; ++ct)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = rbase + i;
                if (row < sg.K) out[(int64_t)row * NP + ct * 16 + l16] = acc[rt][ct][i];
            }
    }
}

// Per-plane data of the gather.
struct GPlane {
    int lo, hi, off, t_lo, t_hi, seg0;
};

// dH(i,j,a,b,c) = dW[seg(c, dx)][(t_hi - dy) * S + ab][uv] with
// n*dx + u = i + a - off, n*dy + v = j + b - off; 0 outside the plane's range.
__global__ void k_dw_gather(const float *__restrict__ dW, const int64_t *__restrict__ dw_off,
                            const GPlane *__restrict__ planes, int64_t M, int n, int nz, int S, int NP,
                            float *__restrict__ dH)
{
    const int64_t total = M * M * n * n * nz;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        int64_t t = e;
        const int i = (int)(t % M); t /= M;
        const int j = (int)(t % M); t /= M;
        const int a = (int)(t % n); t /= n;
        const int b = (int)(t % n); t /= n;
        const int c = (int)t;
        const GPlane pl = planes[c];
        float v = 0.0f;
        if (pl.seg0 >= 0 && i >= pl.lo && i <= pl.hi && j >= pl.lo && j <= pl.hi) {
            const int di = i + a - pl.off, dj = j + b - pl.off;
            const int dx = di >= 0 ? di / n : -((-di + n - 1) / n);
            const int dy = dj >= 0 ? dj / n : -((-dj + n - 1) / n);
            const int u = di - n * dx, w = dj - n * dy;
            if (dx >= pl.t_lo && dx <= pl.t_hi && dy >= pl.t_lo && dy <= pl.t_hi) {
                const int seg = pl.seg0 + (dx - pl.t_lo);
                const int64_t k = (int64_t)(pl.t_hi - dy) * S + (a + n * b);
                v = dW[dw_off[seg] + k * NP + (u + n * w)];
            }
        }
        dH[e] = v;
    }
}

static int grid_of(int64_t total) { return (int)std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 65536)); }

// Batch state kept between the per-frame passes and the PSF-gradient GEMM.
struct GradBuffers {
    float *vols = nullptr, *res = nullptr, *dW = nullptr;
    double *partial = nullptr;
    int64_t *d_dw_off = nullptr;
    DwItem *d_items = nullptr;
    GPlane *d_planes = nullptr;
    ~GradBuffers()
    {
        for (void *q : {(void *)vols, (void *)res, (void *)dW, (void *)partial, (void *)d_dw_off, (void *)d_items,
                        (void *)d_planes})
            if (q) (void)hipFree(q);
    }
};

template <int NT>
static void launch_dw(const GemmPlan &F, const std::vector<DwItem> &items, GradBuffers &gb, const SensorPlan &sp,
                      const sid_psf *p, int nb, hipStream_t st)
{
    hipLaunchKernelGGL((k_dw_gemm<NT>), dim3((unsigned)items.size()), dim3(256), 0, st, F.d_segs, gb.d_items, gb.vols,
                       sp.vol_elems, gb.res, sp.img_elems, nb, (int)sp.nrows, (int)sp.Ly, sp.YP, (int)p->T, (int)p->S,
                       gb.dW, gb.d_dw_off);
}

// mode 0: Poisson likelihood (y = photon counts, residual r = 1 - y/F);
// mode 1: vector-Jacobian product of the forward projector (y = upstream
// gradient images G, r = G; nll unused).
int grad_core(sid_psf *p, const float *V, const float *y, const int64_t dims[4], int mode, double *nll, float *gV,
              float *gH, hipStream_t st)
{
    const int64_t X = dims[0], Y = dims[1], nz = dims[2], nb = dims[3];
    if (p->precision != SID_PREC_FP32) {
        set_error("the likelihood gradient runs in FP32 (set precision fp32)");
        return SID_ERR_STATE;
    }
    // the PSF gradient is the weight gradient of the direct GEMM: run every
    // plane on the direct path for this call
    const int saved_method = p->method;
    p->method = SID_METHOD_DIRECT;
    const int prc = ensure_plan(p, X, Y, nz);
    p->method = saved_method;
    SID_TRY(prc);
    SensorPlan &sp = p->plan;
    GradBuffers gb;
    const int nblk = grid_of(X * Y);
    SID_TRY_HIP(hipMalloc(&gb.partial, sizeof(double) * nblk * nb));
    if (gH) {
        SID_TRY_HIP(hipMalloc(&gb.vols, sizeof(float) * sp.vol_elems * nb));
        SID_TRY_HIP(hipMalloc(&gb.res, sizeof(float) * sp.img_elems * nb));
    }
    const int64_t nvol = X * Y * nz, nimg = X * Y;
    for (int64_t f = 0; f < nb; ++f) {
        if (mode == 0 || gH) SID_TRY(launch_vol_in(p, V + f * nvol, nz, 0, st));
        if (mode == 0) {
            SID_TRY(launch_gemm(p, 0, st));
            SID_TRY(launch_fwd_reduce(p, nullptr, st));
            hipLaunchKernelGGL(k_nll_residual, dim3(nblk), dim3(256), 0, st, sp.img, y + f * nimg, (int)X, (int)Y,
                               (int)p->n, sp.YP, (int)p->T, (int)p->S, gb.partial + f * nblk);
            SID_TRY_HIP(hipGetLastError());
        } else {
            SID_TRY(launch_img_in(p, y + f * nimg, st));
        }
        if (gH) {
            SID_TRY_HIP(hipMemcpyAsync(gb.vols + f * sp.vol_elems, sp.vol, sizeof(float) * sp.vol_elems,
                                       hipMemcpyDeviceToDevice, st));
            SID_TRY_HIP(hipMemcpyAsync(gb.res + f * sp.img_elems, sp.img, sizeof(float) * sp.img_elems,
                                       hipMemcpyDeviceToDevice, st));
        }
        if (gV) {
            SID_TRY(launch_gemm(p, 1, st));
            SID_TRY(launch_bwd_reduce(p, gV + f * X * Y * p->nzH, 0, st));
        }
    }
    if (gH) {
        const GemmPlan &F = sp.fwd;
        std::vector<int64_t> dw_off(F.segs.size());
        int64_t tot = 0;
        std::vector<DwItem> items;
        for (size_t s = 0; s < F.segs.size(); ++s) {
            dw_off[s] = tot;
            tot += (int64_t)F.segs[s].K * p->NP;
            for (int kt = 0; kt * 128 < F.segs[s].K; ++kt) items.push_back(DwItem{(int32_t)s, kt});
        }
        std::vector<GPlane> gp(p->nzH);
        {
            int seg = 0;
            for (int64_t c = 0; c < p->nzH; ++c) {
                const Plane &pl = p->planes[c];
                GPlane g{pl.lo, pl.hi, pl.off_f, pl.tf_lo, pl.tf_hi, -1};
                if (!pl.empty && c < sp.nz) {
                    g.seg0 = seg;
                    seg += pl.tf_hi - pl.tf_lo + 1;
                }
                gp[c] = g;
            }
        }
        SID_TRY_HIP(hipMalloc(&gb.dW, sizeof(float) * std::max<int64_t>(tot, 1)));
        SID_TRY_HIP(hipMalloc(&gb.d_dw_off, sizeof(int64_t) * std::max<size_t>(dw_off.size(), 1)));
        SID_TRY_HIP(hipMalloc(&gb.d_items, sizeof(DwItem) * std::max<size_t>(items.size(), 1)));
        SID_TRY_HIP(hipMalloc(&gb.d_planes, sizeof(GPlane) * gp.size()));
        if (!dw_off.empty())
            SID_TRY_HIP(hipMemcpyAsync(gb.d_dw_off, dw_off.data(), sizeof(int64_t) * dw_off.size(),
                                       hipMemcpyHostToDevice, st));
        if (!items.empty())
            SID_TRY_HIP(hipMemcpyAsync(gb.d_items, items.data(), sizeof(DwItem) * items.size(), hipMemcpyHostToDevice,
                                       st));
        SID_TRY_HIP(hipMemcpyAsync(gb.d_planes, gp.data(), sizeof(GPlane) * gp.size(), hipMemcpyHostToDevice, st));
        if (!items.empty()) {
            switch (p->NP / 16) {
                case 1: launch_dw<1>(F, items, gb, sp, p, (int)nb, st); break;
                case 2: launch_dw<2>(F, items, gb, sp, p, (int)nb, st); break;
                case 4: launch_dw<4>(F, items, gb, sp, p, (int)nb, st); break;
                case 8: launch_dw<8>(F, items, gb, sp, p, (int)nb, st); break;
                case 11: launch_dw<11>(F, items, gb, sp, p, (int)nb, st); break;
                case 15: launch_dw<15>(F, items, gb, sp, p, (int)nb, st); break;
                case 16: launch_dw<16>(F, items, gb, sp, p, (int)nb, st); break;
                default: set_error("unsupported NP"); return SID_ERR_DIMS;
            }
            SID_TRY_HIP(hipGetLastError());
        }
        const int64_t hsz = p->M * p->M * p->n * p->n * p->nzH;
        hipLaunchKernelGGL(k_dw_gather, dim3(grid_of(hsz)), dim3(256), 0, st, gb.dW, gb.d_dw_off, gb.d_planes, p->M,
                           (int)p->n, (int)p->nzH, (int)p->S, (int)p->NP, gH);
        SID_TRY_HIP(hipGetLastError());
    }
    if (mode != 0) {
        SID_TRY_HIP(hipStreamSynchronize(st));
        return SID_OK;
    }
    std::vector<double> part(nblk * nb);
    SID_TRY_HIP(hipMemcpyAsync(part.data(), gb.partial, sizeof(double) * part.size(), hipMemcpyDeviceToHost, st));
    SID_TRY_HIP(hipStreamSynchronize(st));
    double s = 0.0;
    for (double v : part) s += v;  // fixed order: deterministic
    if (nll) *nll = s;
    return SID_OK;
}

// ---------------------------------------------------------------------------
// FP64 back-projection of a sensor of ones: the normalisation B(1) of the
// Richardson-Lucy / EM updates.  backwardProjectGPU_new(H, ones(X, Y)) at voxel
// q of plane c is the sum of the kernel H_{a(q),b(q),c} over the indices
// i = p - q + off_b that land on the sensor, i.e. over a rectangle of the
// kernel's support; each kernel's FP64 summed-area table gives that sum from
// four lookups.  Computed plane by plane (the tables of one plane: n^2 kernels
// of (Mc + 1)^2 doubles).
// ---------------------------------------------------------------------------

// Summed-area table of kernel ab of plane c over its support [lo, hi]^2:
// sat[(j + 1) * W + (i + 1)] = sum of H[lo..lo+i][lo..lo+j], first row / column 0.
__global__ __launch_bounds__(256) void k_b1_sat(const float *__restrict__ H, int64_t M, int n, int c, int lo, int Mc,
                                                double *__restrict__ sat)
{
    const int ab = blockIdx.x, a = ab % n, b = ab / n;
    const int W = Mc + 1;
    const float *K = H + M * M * ((int64_t)a + (int64_t)n * ((int64_t)b + (int64_t)n * c));
    double *T = sat + (int64_t)ab * W * W;
    // rows: running sums along i for every column j
    for (int j = threadIdx.x; j < W; j += blockDim.x) {
        double acc = 0.0;
        T[(int64_t)j * W] = 0.0;
        for (int i = 0; i < Mc; ++i) {
            if (j > 0) acc += (double)K[(int64_t)(lo + i) + M * (lo + j - 1)];
            T[(int64_t)j * W + i + 1] = acc;
        }
    }
    __syncthreads();
    // columns: running sums along j
    for (int i = threadIdx.x; i < W; i += blockDim.x) {
        double acc = 0.0;
        for (int j = 0; j < W; ++j) {
            acc += T[(int64_t)j * W + i];
            T[(int64_t)j * W + i] = acc;
        }
    }
}

// B1[x + X (y + Y c)] = sum of kernel (x mod n, y mod n, c) over i in
// [max(lo, off - x), min(hi, X - 1 - x + off)] x the same for y.
__global__ void k_b1_rect(const double *__restrict__ sat, int X, int Y, int n, int c, int lo, int hi, int off,
                          double *__restrict__ B1)
{
    const int W = hi - lo + 2;
    const int64_t total = (int64_t)X * Y;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(e % X), y = (int)(e / X);
        const int i0 = max(lo, off - x), i1 = min(hi, X - 1 - x + off);
        const int j0 = max(lo, off - y), j1 = min(hi, Y - 1 - y + off);
        double v = 0.0;
        if (i0 <= i1 && j0 <= j1) {
            const double *T = sat + (int64_t)((x % n) + n * (y % n)) * W * W;
            const int u0 = i0 - lo, u1 = i1 - lo + 1, w0 = j0 - lo, w1 = j1 - lo + 1;
            v = T[(int64_t)w1 * W + u1] - T[(int64_t)w0 * W + u1] - T[(int64_t)w1 * W + u0] + T[(int64_t)w0 * W + u0];
        }
        B1[(int64_t)X * Y * c + e] = v;
    }
}

int backward_ones_f64(sid_psf *p, int64_t X, int64_t Y, double *B1, hipStream_t st)
{
    const int n = (int)p->n;
    int64_t tab = 1;
    for (const Plane &pl : p->planes)
        if (!pl.empty) tab = std::max<int64_t>(tab, (int64_t)(pl.hi - pl.lo + 2) * (pl.hi - pl.lo + 2));
    double *sat = nullptr;
    SID_TRY_HIP(hipMalloc(&sat, sizeof(double) * tab * n * n));
    int rc = SID_OK;
    for (int64_t c = 0; c < p->nzH && rc == SID_OK; ++c) {
        const Plane &pl = p->planes[c];
        if (pl.empty) {
            hipLaunchKernelGGL(k_b1_rect, dim3(grid_of(X * Y)), dim3(256), 0, st, (const double *)nullptr, (int)X,
                               (int)Y, n, (int)c, 1, 0, 0, B1);   // empty range: zeros
        } else {
            const int Mc = pl.hi - pl.lo + 1;
            hipLaunchKernelGGL(k_b1_sat, dim3((unsigned)(n * n)), dim3(256), 0, st, p->H, p->M, n, (int)c, pl.lo,
                               Mc, sat);
            hipLaunchKernelGGL(k_b1_rect, dim3(grid_of(X * Y)), dim3(256), 0, st, (const double *)sat, (int)X,
                               (int)Y, n, (int)c, pl.lo, pl.hi, pl.off_b, B1);
        }
        if (hipGetLastError() != hipSuccess) {
            set_error("B(1) kernels failed to launch");
            rc = SID_ERR_HIP;
        }
    }
    const hipError_t e = hipStreamSynchronize(st);
    (void)hipFree(sat);
    if (rc == SID_OK && e != hipSuccess) {
        set_error("B(1): %s", hipGetErrorString(e));
        rc = SID_ERR_HIP;
    }
    return rc;
}

}  // namespace sid
