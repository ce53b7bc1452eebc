// The kernel option of the batched library (src/generate_LFM_library_GPU.m:
// 54-56,95-101): with options.ker_param set, each seed's sub-volume — the
// average of the listed reconstructions on the seed's clipped box, zero
// elsewhere in an image_size(1) x image_size(2) x (z1 - z0 + 1) volume (:70-85)
// — is convolved with the kernel (convn(Volume, kern, 'same') on the GPU in
// single precision) before the forward projection with planes z0..z1 of H.
//
// convn 'same' of a volume that is zero outside the box is nonzero only on the
// box grown by the kernel's reach, out[x] = sum_i V[x + ox - i] k[i] with
// ox = floor(kx / 2): x in [x0 - ox, x1 + kx - 1 - ox] (and likewise in y),
// clipped to the sensor (the 'same' crop of the full-lateral volume); in z the
// output keeps the box's planes (the 'same' crop of the z0..z1 volume).  So a
// seed with a kernel is a seed on the grown box, whose values come from one
// convn over an atlas: the seeds' boxes side by side along x, each in a slot
// of (bw + 2 (kx - 1)) x (bh + 2 (ky - 1)) x (bz + 2 (kz - 1)) voxels with its
// values at offset (kx - 1, ky - 1, kz - 1) and zeros around them.  A slot's
// zero margin is at least the kernel's reach on both sides, so no seed's
// values reach another seed's output region, and the output region of each
// seed is read back from the atlas (FP32 throughout, the library's own
// arithmetic).
#include <algorithm>
#include <vector>

#include "sid_internal.h"

namespace sid {

// Per seed: values offsets, atlas slot, box and grown-box sizes.
struct SeedConv {
    int64_t val_off, out_off, ax;   // input values, output values, slot start along x
    int bw, bh, bz;                 // box sizes
    int gdx, gdy;                   // grown box start minus box start (<= 0)
    int gw, gh;                     // grown box sizes
};

__global__ __launch_bounds__(256) void k_seed_scatter(const float *__restrict__ vals, const SeedConv *__restrict__ sc,
                                                      int64_t AX, int64_t AY, int kx, int ky, int kz,
                                                      float *__restrict__ atlas)
{
    const SeedConv s = sc[blockIdx.x];
    const int64_t nel = (int64_t)s.bw * s.bh * s.bz;
    for (int64_t e = threadIdx.x; e < nel; e += blockDim.x) {
        const int i = (int)(e % s.bw), j = (int)((e / s.bw) % s.bh), l = (int)(e / ((int64_t)s.bw * s.bh));
        atlas[(s.ax + kx - 1 + i) + AX * ((int64_t)(ky - 1 + j) + AY * (int64_t)(kz - 1 + l))] = vals[s.val_off + e];
    }
}

__global__ __launch_bounds__(256) void k_seed_gather(const float *__restrict__ atlas, const SeedConv *__restrict__ sc,
                                                     int64_t AX, int64_t AY, int kx, int ky, int kz,
                                                     float *__restrict__ out)
{
    const SeedConv s = sc[blockIdx.x];
    const int64_t nel = (int64_t)s.gw * s.gh * s.bz;
    for (int64_t e = threadIdx.x; e < nel; e += blockDim.x) {
        const int u = (int)(e % s.gw), v = (int)((e / s.gw) % s.gh), l = (int)(e / ((int64_t)s.gw * s.gh));
        const int64_t x = s.ax + kx - 1 + s.gdx + u, y = ky - 1 + s.gdy + v;
        out[s.out_off + e] = atlas[x + AX * (y + AY * (int64_t)(kz - 1 + l))];
    }
}

int library_kernel_boxes(int64_t ns, const int64_t *boxes, int64_t X, int64_t Y, const int kd[3], int64_t *out,
                         int64_t *voff_out)
{
    const int ox = kd[0] / 2, oy = kd[1] / 2;
    if (voff_out) voff_out[0] = 0;
    for (int64_t s = 0; s < ns; ++s) {
        const int64_t *b = boxes + 6 * s;
        if (b[0] < 1 || b[1] > X || b[0] > b[1] || b[2] < 1 || b[3] > Y || b[2] > b[3] || b[4] < 1 || b[4] > b[5]) {
            set_error("seed %lld: box outside the %lld x %lld sensor or empty", (long long)s, (long long)X,
                      (long long)Y);
            return SID_ERR_DIMS;
        }
        int64_t *o = out + 6 * s;
        o[0] = std::max<int64_t>(1, b[0] - ox);
        o[1] = std::min<int64_t>(X, b[1] + kd[0] - 1 - ox);
        o[2] = std::max<int64_t>(1, b[2] - oy);
        o[3] = std::min<int64_t>(Y, b[3] + kd[1] - 1 - oy);
        o[4] = b[4];
        o[5] = b[5];
        if (voff_out) voff_out[s + 1] = voff_out[s] + (o[1] - o[0] + 1) * (o[3] - o[2] + 1) * (o[5] - o[4] + 1);
    }
    return SID_OK;
}

int library_convn_seeds(int64_t ns, const int64_t *boxes, const float *vals, const int64_t *voff, int64_t X,
                        int64_t Y, const float *k_dev, const int kd[3], float *out, hipStream_t st)
{
    if (kd[0] < 1 || kd[1] < 1 || kd[2] < 1 || kd[0] > kConvnMaxXY || kd[1] > kConvnMaxXY || kd[2] > kConvnMaxZ) {
        set_error("library kernel %d x %d x %d outside 1..%d x 1..%d x 1..%d", kd[0], kd[1], kd[2], kConvnMaxXY,
                  kConvnMaxXY, kConvnMaxZ);
        return SID_ERR_DIMS;
    }
    if (ns <= 0) return SID_OK;
    std::vector<int64_t> grown(6 * ns), goff(ns + 1);
    SID_TRY(library_kernel_boxes(ns, boxes, X, Y, kd, grown.data(), goff.data()));
    std::vector<SeedConv> sc(ns);
    for (int64_t s = 0; s < ns; ++s) {
        const int64_t *b = boxes + 6 * s, *g = grown.data() + 6 * s;
        SeedConv &c = sc[s];
        c.bw = (int)(b[1] - b[0] + 1);
        c.bh = (int)(b[3] - b[2] + 1);
        c.bz = (int)(b[5] - b[4] + 1);
        if (voff[s + 1] - voff[s] != (int64_t)c.bw * c.bh * c.bz) {
            set_error("seed %lld: %lld values for a %d x %d x %d box", (long long)s,
                      (long long)(voff[s + 1] - voff[s]), c.bw, c.bh, c.bz);
            return SID_ERR_DIMS;
        }
        c.val_off = voff[s];
        c.out_off = goff[s];
        c.gdx = (int)(g[0] - b[0]);
        c.gdy = (int)(g[2] - b[2]);
        c.gw = (int)(g[1] - g[0] + 1);
        c.gh = (int)(g[3] - g[2] + 1);
    }
    // batches of seeds whose atlas stays within the convn's x range and 512 MB
    constexpr int64_t kMaxAtlas = (int64_t)128 << 20;   // floats
    SeedConv *d_sc = nullptr;
    float *atlas = nullptr, *conv = nullptr;
    int64_t cap = 0;
    int rc = SID_OK;
    if (hipMalloc(&d_sc, sizeof(SeedConv) * ns) != hipSuccess) {
        set_error("hipMalloc of the seed table failed");
        return SID_ERR_OOM;
    }
    for (int64_t s0 = 0; s0 < ns && rc == SID_OK;) {
        int64_t AX = 0, AY = 0, AZ = 0, s1 = s0;
        while (s1 < ns) {
            const SeedConv &c = sc[s1];
            const int64_t w = c.bw + 2 * (kd[0] - 1);
            const int64_t nAY = std::max<int64_t>(AY, c.bh + 2 * (kd[1] - 1));
            const int64_t nAZ = std::max<int64_t>(AZ, c.bz + 2 * (kd[2] - 1));
            if (s1 > s0 && ((AX + w) * nAY * nAZ > kMaxAtlas || AX + w > (1 << 24))) break;
            sc[s1].ax = AX;
            AX += w;
            AY = nAY;
            AZ = nAZ;
            ++s1;
        }
        const int64_t nvox = AX * AY * AZ;
        if (nvox > cap) {
            if (atlas) (void)hipFree(atlas);
            if (conv) (void)hipFree(conv);
            atlas = conv = nullptr;
            if (hipMalloc(&atlas, sizeof(float) * nvox) != hipSuccess ||
                hipMalloc(&conv, sizeof(float) * nvox) != hipSuccess) {
                set_error("hipMalloc of the %lld-voxel seed atlas failed", (long long)nvox);
                rc = SID_ERR_OOM;
                break;
            }
            cap = nvox;
        }
        if (hipMemcpyAsync(d_sc + s0, sc.data() + s0, sizeof(SeedConv) * (s1 - s0), hipMemcpyHostToDevice, st) !=
                hipSuccess ||
            hipMemsetAsync(atlas, 0, sizeof(float) * nvox, st) != hipSuccess) {
            set_error("seed atlas set-up failed");
            rc = SID_ERR_HIP;
            break;
        }
        hipLaunchKernelGGL(k_seed_scatter, dim3((unsigned)(s1 - s0)), dim3(256), 0, st, vals, d_sc + s0, AX, AY, kd[0],
                           kd[1], kd[2], atlas);
        if (hipGetLastError() != hipSuccess) {
            set_error("k_seed_scatter launch failed");
            rc = SID_ERR_HIP;
            break;
        }
        const int64_t adims[3] = {AX, AY, AZ};
        rc = convn_same(atlas, adims, k_dev, kd, conv, st);
        if (rc != SID_OK) break;
        hipLaunchKernelGGL(k_seed_gather, dim3((unsigned)(s1 - s0)), dim3(256), 0, st, conv, d_sc + s0, AX, AY, kd[0],
                           kd[1], kd[2], out);
        if (hipGetLastError() != hipSuccess) {
            set_error("k_seed_gather launch failed");
            rc = SID_ERR_HIP;
            break;
        }
        // the next batch reuses the buffers: wait for this one
        if (hipStreamSynchronize(st) != hipSuccess) {
            set_error("seed convolution failed");
            rc = SID_ERR_HIP;
        }
        s0 = s1;
    }
    (void)hipStreamSynchronize(st);
    if (atlas) (void)hipFree(atlas);
    if (conv) (void)hipFree(conv);
    (void)hipFree(d_sc);
    return rc;
}

}  // namespace sid
