// Batched light-field library projection: the per-seed forward projections
// of src/generate_LFM_library_GPU.m:63-108 (each seed a small sub-volume,
// projected with the planes of H it spans, :87-88,95) for thousands of seeds
// in one persistent launch, written as footprint-cropped sensor patches whose
// rectangles follow src/project_forward_patched.m:8-20,54-69.
//
// The seeds' sub-volumes are mostly empty at full-sensor scale, so instead of
// the output-stationary Toeplitz GEMM of sid_direct.hip the work is organised
// tap-major: for plane c and lenslet tap D = (dx, dy)
//     C[(s, Q), uv] = sum_ab A[(s, Q), ab] * W_c[D][ab][uv]
// where the rows (s, Q) are the (seed, input lenslet) pairs active on plane c
// across all seeds (the B operand W_c[D] is shared by every seed), and each
// row of C is added to seed s's accumulator at output lenslet Q + D
// (FP32 atomics; summation order is not fixed).  A persistent grid pulls
// (plane, row tile, dx) items from a queue.  The whole sequence (fills, row
// build, GEMM, patch extraction) is captured once in a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "sid_internal.h"

namespace sid {

constexpr int LIB_MT = 64;     // rows per item (4 waves x 16)
constexpr int LIB_KROW = 256;  // A row stride: S entries padded with zeros to 8 chunks of 32

struct LibRow {                // one (seed, input lenslet, plane) row
    int32_t seed, Qx, Qy, plane;
    int64_t acc_base;          // accumulator offset of output lenslet Q (tap 0)
    int32_t wy, pad;           // accumulator row width (lenslets) of the seed
};

struct LibItem {
    int32_t row0, nrows, plane, dx;
    int64_t w_off;             // weights of (plane, dx): chunked [K/32][NP][32]
    int32_t t_lo, t_hi;
    int32_t live;              // offset (words) of the (plane, dx) chunk table: bit 8 (dy - t_lo) + kc / 32
    int32_t pad;               //   set when chunk kc of tap dy has a nonzero weight
};

struct LibSeed {
    int32_t x0, x1, y0, y1, z0, z1;      // box, 0-based inclusive
    int32_t sx0, sx1, sy0, sy1;          // output rectangle, 0-based inclusive
    int32_t ax0, ay0, wx, wy;            // accumulator origin (lenslets) and size
    int64_t val_off, acc_off, out_off;
};

__global__ void k_fill_f32(float *__restrict__ p, int64_t n, float v)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        p[e] = v;
}

__global__ void k_fill_i32(int *__restrict__ p, int n, int v)
{
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) p[e] = v;
}

// A[row][ab] from the seeds' column-major sub-volumes; zero outside the box
// and for ab >= n*n.
__global__ void k_lib_rows(const LibRow *__restrict__ rows, int nrows, const LibSeed *__restrict__ seeds,
                           const float *__restrict__ vals, int n, float *__restrict__ A)
{
    const int64_t total = (int64_t)nrows * LIB_KROW;
    const int nn2 = n * n;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(e / LIB_KROW), ab = (int)(e - (int64_t)r * LIB_KROW);
        float v = 0.0f;
        if (ab < nn2) {
            const LibRow rw = rows[r];
            const LibSeed sd = seeds[rw.seed];
            const int x = rw.Qx * n + ab % n, y = rw.Qy * n + ab / n;
            if (x >= sd.x0 && x <= sd.x1 && y >= sd.y0 && y <= sd.y1) {
                const int bx = sd.x1 - sd.x0 + 1, by = sd.y1 - sd.y0 + 1;
                v = vals[sd.val_off + (x - sd.x0) + (int64_t)bx * ((y - sd.y0) + (int64_t)by * (rw.plane - sd.z0))];
            }
        }
        A[e] = v;
    }
}

// Persistent tap-major GEMM.  256 threads = 4 waves of 16 rows; NT column
// tiles of 16; v_mfma_f32_16x16x4_f32 with the same lane/K mapping as the
// FP32 projector (lane group g supplies k = 8g + s at step s of a 32-chunk).
template <int NT>
__global__ __launch_bounds__(256) void k_lib_gemm(const LibItem *__restrict__ items, int n_items,
                                                  int *__restrict__ counter, const LibRow *__restrict__ rows,
                                                  const float *__restrict__ A, const float *__restrict__ W, int S,
                                                  int nn2, const uint32_t *__restrict__ live_tab,
                                                  float *__restrict__ acc_buf)
{
    constexpr int NP = NT * 16, KC = 32, LDK = KC + 4;
    __shared__ float Bs[NP * LDK];
    __shared__ int64_t row_base[LIB_MT];
    __shared__ int row_wy[LIB_MT];
    __shared__ int cur;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, g = lane >> 4;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) cur = atomicAdd(counter, 1);
        __syncthreads();
        const int idx = cur;
        if (idx >= n_items) break;   // every wave of the block sees the same value
        const LibItem it = items[idx];
        if (threadIdx.x < LIB_MT) {
            const int r = it.row0 + (int)threadIdx.x;
            if ((int)threadIdx.x < it.nrows) {
                row_base[threadIdx.x] = rows[r].acc_base;
                row_wy[threadIdx.x] = rows[r].wy;
            } else {
                row_base[threadIdx.x] = -1;
                row_wy[threadIdx.x] = 0;
            }
        }
        const int arow = it.row0 + min(wave * 16 + l16, it.nrows - 1);
        const float *Arow = A + (int64_t)arow * LIB_KROW;
        const float *Wp = W + it.w_off;
        for (int dy = it.t_lo; dy <= it.t_hi; ++dy) {
            // chunks of this tap with a nonzero weight (uniform); a tap without any adds nothing
            const int tb = 8 * (dy - it.t_lo);
            const uint32_t live = (live_tab[it.live + (tb >> 5)] >> (tb & 31)) & 0xffu;
            if (live == 0) continue;
            const int kb = (it.t_hi - dy) * S;
            f4 acc[NT];
#pragma unroll
            for (int ct = 0; ct < NT; ++ct) acc[ct] = f4{0.f, 0.f, 0.f, 0.f};
            // the weights of a live chunk: loaded into registers one chunk ahead
            // (their loads land under the previous chunk's MFMAs), then staged
            constexpr int WPT = (NP * (KC / 4) + 255) / 256;   // float4 per thread and chunk
            float4 wreg[WPT];
            auto wload = [&](int kc) {
#pragma unroll
                for (int q = 0; q < WPT; ++q) {
                    const int e = (int)threadIdx.x + 256 * q;
                    const int col = e / (KC / 4), kq = e - col * (KC / 4);
                    const int kl = kc + 4 * kq;  // local k within the dy block
                    float4 w = {0.f, 0.f, 0.f, 0.f};
                    if (e < NP * (KC / 4) && kl < S) {
                        const int k = kb + kl;   // multiple of 4: one float4 never crosses a chunk
                        w = *(const float4 *)(Wp + (int64_t)(k / KC) * NP * KC + (int64_t)col * KC + (k % KC));
                    }
                    wreg[q] = w;
                }
            };
            int kc = 0;
            while (kc < LIB_KROW && !((live >> (kc / KC)) & 1u)) kc += KC;
            wload(kc);
            for (; kc < LIB_KROW;) {
                int kn = kc + KC;   // the next live chunk of this tap
                while (kn < LIB_KROW && !((live >> (kn / KC)) & 1u)) kn += KC;
                const float4 a0 = *(const float4 *)(Arow + kc + 8 * g);
                const float4 a1 = *(const float4 *)(Arow + kc + 8 * g + 4);
                __syncthreads();
#pragma unroll
                for (int q = 0; q < WPT; ++q) {
                    const int e = (int)threadIdx.x + 256 * q;
                    if (e < NP * (KC / 4)) {
                        const int col = e / (KC / 4), kq = e - col * (KC / 4);
                        *(float4 *)(Bs + col * LDK + 4 * kq) = wreg[q];
                    }
                }
                if (kn < LIB_KROW) wload(kn);
                __syncthreads();
#pragma unroll
                for (int ct = 0; ct < NT; ++ct) {
                    const float *pb = Bs + (ct * 16 + l16) * LDK + 8 * g;
                    const float4 b0 = *(const float4 *)pb, b1 = *(const float4 *)(pb + 4);
                    f4 c = acc[ct];
                    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a0.x, b0.x, c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a0.y, b0.y, c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a0.z, b0.z, c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a0.w, b0.w, c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a1.x, b1.x, c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a1.y, b1.y, c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a1.z, b1.z, c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a1.w, b1.w, c, 0, 0, 0);
                    acc[ct] = c;
                }
                kc = kn;
            }
            // scatter-add: lane holds C[4g + i][ct*16 + l16]
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int lr = wave * 16 + 4 * g + i;
                const int64_t base = row_base[lr];
                if (base < 0) continue;
                const int64_t o = base + ((int64_t)it.dx * row_wy[lr] + dy) * S;
#pragma unroll
                for (int ct = 0; ct < NT; ++ct) {
                    const int col = ct * 16 + l16;
                    if (col < nn2) atomicAdd(acc_buf + o + col, acc[ct][i]);
                }
            }
        }
    }
}

// Accumulators (lenslet-major) -> column-major sensor patches.
__global__ void k_lib_out(const LibSeed *__restrict__ seeds, int n_seeds, const float *__restrict__ acc_buf, int n,
                          int S, float *__restrict__ out, const int64_t *__restrict__ pix_off)
{
    // one block row per seed (gridDim.y), blocks over the patch pixels
    const int s = blockIdx.y;
    if (s >= n_seeds) return;
    const LibSeed sd = seeds[s];
    const int pw = sd.sx1 - sd.sx0 + 1, ph = sd.sy1 - sd.sy0 + 1;
    const int64_t total = (int64_t)pw * ph;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int x = sd.sx0 + (int)(e % pw), y = sd.sy0 + (int)(e / pw);
        const int Px = x / n - sd.ax0, Py = y / n - sd.ay0;
        float v = 0.0f;
        if (Px >= 0 && Px < sd.wx && Py >= 0 && Py < sd.wy)
            v = acc_buf[sd.acc_off + ((int64_t)Px * sd.wy + Py) * S + (x % n) + n * (y % n)];
        out[pix_off[s] + e] = v;
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

static inline int64_t cdiv_i(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Output rectangle of src/project_forward_patched.m:8-20,54-59 for a seed
// box (1-based inclusive px_min..px_max), kernel side M, sensor X x Y;
// 1-based inclusive result.
void patched_rect(int64_t n, int64_t M, int64_t X, int64_t Y, const int64_t box[6], int64_t rect[4])
{
    for (int d = 0; d < 2; ++d) {
        const int64_t lo = box[2 * d], hi = box[2 * d + 1], size = d == 0 ? X : Y;
        const int64_t off_ml = (lo - 1) / n;             // floor((px_min - 1) / n), px_min >= 1
        const int64_t off_px = off_ml * n;
        int64_t n_ml = cdiv_i(hi, n) - off_ml;           // ceil(px_max / n) - off_ml
        if (n_ml % 2 == 0) n_ml += 1;                    // always odd
        const int64_t pj = n_ml * n + M - 1;             // conv2 'full' size
        const int64_t center = off_px + cdiv_i(n_ml * n, 2);  // ceil((n_ml / 2) * n)
        rect[2 * d] = std::max<int64_t>(1, center - pj / 2);
        rect[2 * d + 1] = std::min<int64_t>(size, center + pj / 2);
    }
}

struct LibPlan {
    // key
    std::vector<int64_t> boxes, val_off;
    int64_t X = 0, Y = 0;
    const float *vals = nullptr;
    float *out = nullptr;
    // device state
    LibRow *d_rows = nullptr;
    LibItem *d_items = nullptr;
    LibSeed *d_seeds = nullptr;
    int64_t *d_pix_off = nullptr;
    int *d_counter = nullptr;
    uint32_t *d_live = nullptr;
    float *d_A = nullptr, *d_acc = nullptr;
    int64_t n_rows = 0, n_items = 0, acc_elems = 0, n_seeds = 0;
    int grid = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int64_t macs = 0;
    ~LibPlan()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void *q : {(void *)d_rows, (void *)d_items, (void *)d_seeds, (void *)d_pix_off, (void *)d_counter,
                        (void *)d_live, (void *)d_A, (void *)d_acc})
            if (q) (void)hipFree(q);
    }
};

void lib_plan_free(sid_psf *p)
{
    delete (LibPlan *)p->lib;
    p->lib = nullptr;
}

static int enqueue(sid_psf *p, LibPlan &L, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, L.d_counter, 1, 0);
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)std::min<int64_t>(cdiv_i(L.acc_elems, 256), 65536)), dim3(256), 0,
                       st, L.d_acc, L.acc_elems, 0.0f);
    if (L.n_rows > 0)
        hipLaunchKernelGGL(k_lib_rows, dim3((unsigned)std::min<int64_t>(cdiv_i(L.n_rows * LIB_KROW, 256), 65536)),
                           dim3(256), 0, st, L.d_rows, (int)L.n_rows, L.d_seeds, L.vals, (int)p->n, L.d_A);
    if (L.n_items > 0) {
        const int nn2 = (int)(p->n * p->n);
        switch (p->NP / 16) {
#define SID_LIB_CASE(NTV)                                                                                        \
    case NTV:                                                                                                    \
        hipLaunchKernelGGL((k_lib_gemm<NTV>), dim3(L.grid), dim3(256), 0, st, L.d_items, (int)L.n_items,         \
                           L.d_counter, L.d_rows, L.d_A, p->wf, (int)p->S, nn2, L.d_live, L.d_acc);              \
        break;
            SID_LIB_CASE(1) SID_LIB_CASE(2) SID_LIB_CASE(4) SID_LIB_CASE(8) SID_LIB_CASE(11) SID_LIB_CASE(15)
            SID_LIB_CASE(16)
#undef SID_LIB_CASE
            default:
                set_error("unsupported NP");
                return SID_ERR_DIMS;
        }
    }
    hipLaunchKernelGGL(k_lib_out, dim3(64, (unsigned)L.n_seeds), dim3(256), 0, st, L.d_seeds, (int)L.n_seeds, L.d_acc,
                       (int)p->n, (int)p->S, L.out, L.d_pix_off);
    return SID_OK;
}

int library_forward(sid_psf *p, int64_t ns, const int64_t *boxes, const float *vals, const int64_t *val_off, int64_t X,
                    int64_t Y, float *out, hipStream_t st, int use_graph)
{
    const int64_t n = p->n, S = p->S, NP = p->NP;
    LibPlan *L = (LibPlan *)p->lib;
    const bool same = L && L->X == X && L->Y == Y && L->vals == vals && L->out == out && L->n_seeds == ns &&
                      L->boxes.size() == (size_t)ns * 6 &&
                      std::memcmp(L->boxes.data(), boxes, sizeof(int64_t) * 6 * ns) == 0 &&
                      std::memcmp(L->val_off.data(), val_off, sizeof(int64_t) * ns) == 0;
    if (!same) {
        lib_plan_free(p);
        L = new LibPlan();
        p->lib = L;
        L->boxes.assign(boxes, boxes + 6 * ns);
        L->val_off.assign(val_off, val_off + ns);
        L->X = X; L->Y = Y; L->vals = vals; L->out = out; L->n_seeds = ns;
        std::vector<LibSeed> seeds(ns);
        std::vector<int64_t> pix_off(ns + 1, 0);
        int64_t acc = 0;
        for (int64_t s = 0; s < ns; ++s) {
            const int64_t *b = boxes + 6 * s;
            int64_t rect[4];
            patched_rect(n, p->M, X, Y, b, rect);
            LibSeed &sd = seeds[s];
            sd.x0 = (int)b[0] - 1; sd.x1 = (int)b[1] - 1; sd.y0 = (int)b[2] - 1; sd.y1 = (int)b[3] - 1;
            sd.z0 = (int)b[4] - 1; sd.z1 = (int)b[5] - 1;
            sd.sx0 = (int)rect[0] - 1; sd.sx1 = (int)rect[1] - 1; sd.sy0 = (int)rect[2] - 1; sd.sy1 = (int)rect[3] - 1;
            const int qx0 = sd.x0 / (int)n, qx1 = sd.x1 / (int)n, qy0 = sd.y0 / (int)n, qy1 = sd.y1 / (int)n;
            // the accumulator spans the taps of the seed's own planes only (output
            // lenslet Q + D, D in [tf_lo, tf_hi] of a plane), not the widest plane's
            int t_lo = 0, t_hi = 0;
            for (int c = std::max(sd.z0, 0); c <= sd.z1 && c < (int)p->nzH; ++c)
                if (!p->planes[c].empty) {
                    t_lo = std::min(t_lo, p->planes[c].tf_lo);
                    t_hi = std::max(t_hi, p->planes[c].tf_hi);
                }
            sd.ax0 = qx0 + t_lo; sd.ay0 = qy0 + t_lo;
            sd.wx = qx1 - qx0 + 1 + t_hi - t_lo; sd.wy = qy1 - qy0 + 1 + t_hi - t_lo;
            sd.val_off = val_off[s];
            sd.acc_off = acc;
            acc += (int64_t)sd.wx * sd.wy * S;
            pix_off[s + 1] = pix_off[s] + (int64_t)(sd.sx1 - sd.sx0 + 1) * (sd.sy1 - sd.sy0 + 1);
        }
        L->acc_elems = std::max<int64_t>(acc, 1);
        // rows grouped by plane
        std::vector<LibRow> rows;
        std::vector<LibItem> items;
        // chunk tables per (plane, dx) from the FP32 zero-chunk flags: local chunk kc of tap dy
        // covers K rows (t_hi - dy) S + [32 kc, min(32 kc + 32, S)) of the tap row's weights
        std::vector<uint32_t> live;
        auto live_table = [&](const Plane &pl, int dx) -> int32_t {
            const int ntap = pl.tf_hi - pl.tf_lo + 1;
            const int32_t at = (int32_t)live.size();
            live.resize(live.size() + (8 * ntap + 31) / 32, 0);
            const int64_t c0 = (pl.wf_off + (int64_t)(dx - pl.tf_lo) * pl.Kf * NP) / (NP * 32);
            for (int dy = pl.tf_lo; dy <= pl.tf_hi; ++dy)
                for (int kc = 0; kc < LIB_KROW / 32 && 32 * kc < S; ++kc) {
                    const int64_t k0 = (int64_t)(pl.tf_hi - dy) * S + 32 * kc;
                    const int64_t k1 = (int64_t)(pl.tf_hi - dy) * S + std::min<int64_t>(32 * kc + 32, S);
                    bool any = false;
                    for (int64_t j = k0 / 32; j <= (k1 - 1) / 32 && !any; ++j)
                        any = c0 + j >= (int64_t)p->cflag_f.size() || p->cflag_f[c0 + j];   // unknown: computed
                    const int b = 8 * (dy - pl.tf_lo) + kc;
                    if (any) live[at + b / 32] |= 1u << (b % 32);
                }
            return at;
        };
        for (int64_t c = 0; c < p->nzH; ++c) {
            const Plane &pl = p->planes[c];
            const int64_t r0 = (int64_t)rows.size();
            for (int64_t s = 0; s < ns; ++s) {
                const LibSeed &sd = seeds[s];
                if (c < sd.z0 || c > sd.z1) continue;
                for (int qx = sd.x0 / (int)n; qx <= sd.x1 / (int)n; ++qx)
                    for (int qy = sd.y0 / (int)n; qy <= sd.y1 / (int)n; ++qy) {
                        LibRow rw;
                        rw.seed = (int)s; rw.Qx = qx; rw.Qy = qy; rw.plane = (int)c;
                        rw.wy = sd.wy; rw.pad = 0;
                        rw.acc_base = sd.acc_off + ((int64_t)(qx - sd.ax0) * sd.wy + (qy - sd.ay0)) * S;
                        rows.push_back(rw);
                    }
            }
            const int64_t r1 = (int64_t)rows.size();
            if (pl.empty || r1 == r0) continue;
            std::vector<int32_t> live_at;
            for (int dx = pl.tf_lo; dx <= pl.tf_hi; ++dx) live_at.push_back(live_table(pl, dx));
            for (int64_t r = r0; r < r1; r += LIB_MT)
                for (int dx = pl.tf_lo; dx <= pl.tf_hi; ++dx) {
                    LibItem it;
                    it.row0 = (int)r; it.nrows = (int)std::min<int64_t>(LIB_MT, r1 - r);
                    it.plane = (int)c; it.dx = dx;
                    it.w_off = pl.wf_off + (int64_t)(dx - pl.tf_lo) * pl.Kf * NP;
                    it.t_lo = pl.tf_lo; it.t_hi = pl.tf_hi;
                    it.live = live_at[dx - pl.tf_lo];
                    it.pad = 0;
                    items.push_back(it);
                    int64_t nlive = 0;   // chunks the GEMM runs for this item
                    for (int w = 0; w < (8 * (pl.tf_hi - pl.tf_lo + 1) + 31) / 32; ++w)
                        nlive += __builtin_popcount(live[it.live + w]);
                    L->macs += (int64_t)LIB_MT * 32 * NP * nlive;
                }
        }
        L->n_rows = (int64_t)rows.size();
        L->n_items = (int64_t)items.size();
        int cus = 256;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, p->device) == hipSuccess) cus = prop.multiProcessorCount;
        L->grid = (int)std::max<int64_t>(1, std::min<int64_t>(L->n_items, 2 * cus));
        SID_TRY_HIP(hipMalloc(&L->d_rows, sizeof(LibRow) * std::max<int64_t>(L->n_rows, 1)));
        SID_TRY_HIP(hipMalloc(&L->d_items, sizeof(LibItem) * std::max<int64_t>(L->n_items, 1)));
        SID_TRY_HIP(hipMalloc(&L->d_seeds, sizeof(LibSeed) * std::max<int64_t>(ns, 1)));
        SID_TRY_HIP(hipMalloc(&L->d_pix_off, sizeof(int64_t) * (ns + 1)));
        SID_TRY_HIP(hipMalloc(&L->d_counter, sizeof(int)));
        SID_TRY_HIP(hipMalloc(&L->d_live, sizeof(uint32_t) * std::max<size_t>(live.size(), 1)));
        if (!live.empty())
            SID_TRY_HIP(hipMemcpy(L->d_live, live.data(), sizeof(uint32_t) * live.size(), hipMemcpyHostToDevice));
        SID_TRY_HIP(hipMalloc(&L->d_A, sizeof(float) * std::max<int64_t>(L->n_rows, 1) * LIB_KROW));
        SID_TRY_HIP(hipMalloc(&L->d_acc, sizeof(float) * L->acc_elems));
        if (L->n_rows)
            SID_TRY_HIP(hipMemcpy(L->d_rows, rows.data(), sizeof(LibRow) * L->n_rows, hipMemcpyHostToDevice));
        if (L->n_items)
            SID_TRY_HIP(hipMemcpy(L->d_items, items.data(), sizeof(LibItem) * L->n_items, hipMemcpyHostToDevice));
        if (ns) SID_TRY_HIP(hipMemcpy(L->d_seeds, seeds.data(), sizeof(LibSeed) * ns, hipMemcpyHostToDevice));
        SID_TRY_HIP(hipMemcpy(L->d_pix_off, pix_off.data(), sizeof(int64_t) * (ns + 1), hipMemcpyHostToDevice));
        if (use_graph) {
            // capture on a private stream, then replay on the caller's stream
            hipStream_t cs;
            SID_TRY_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
            SID_TRY_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue(p, *L, cs);
            hipGraph_t gr = nullptr;
            const hipError_t ec = hipStreamEndCapture(cs, &gr);
            (void)hipStreamDestroy(cs);
            if (rc != SID_OK) return rc;
            if (ec != hipSuccess) {
                set_error("hipStreamEndCapture failed: %s", hipGetErrorString(ec));
                return SID_ERR_HIP;
            }
            L->graph = gr;
            SID_TRY_HIP(hipGraphInstantiate(&L->exec, L->graph, nullptr, nullptr, 0));
        }
    }
    if (use_graph && L->exec) {
        SID_TRY_HIP(hipGraphLaunch(L->exec, st));
    } else {
        SID_TRY(enqueue(p, *L, st));
        SID_TRY_HIP(hipGetLastError());
    }
    SID_TRY_HIP(hipStreamSynchronize(st));
    return SID_OK;
}

int library_stats(const sid_psf *p, int64_t st[4])
{
    const LibPlan *L = (const LibPlan *)p->lib;
    st[0] = L ? L->n_rows : 0;
    st[1] = L ? L->n_items : 0;
    st[2] = L ? L->macs : 0;
    st[3] = L ? L->acc_elems : 0;
    return SID_OK;
}

}  // namespace sid
