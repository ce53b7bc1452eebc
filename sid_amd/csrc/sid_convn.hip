// convn(V, k, 'same') on column-major device volumes: the neuron-kernel
// wrapping of the projectors in reconstruct_S.m:122-128,184-190 and
// generate_LFM_library_GPU.m:95-101 (forwardFUN = F(convn(X, k, 'same')),
// backwardFUN = convn(B(y), k, 'same')).
//
// MATLAB's 'same' window of the full convolution starts at floor(size(k)/2):
//   out[x, y, z] = sum_{i,j,l} V[x + ox - i, y + oy - j, z + oz - l] k[i, j, l],
//   (ox, oy, oz) = floor((kx, ky, kz) / 2), V = 0 outside the volume.
//
// One block computes a 64 x 16 (x, y) output tile of one plane in FP32.  For
// every kernel slice l the (64 + kx - 1) x (16 + ky - 1) input tile of plane
// z + oz - l is staged in LDS; a thread owns 4 consecutive x outputs and walks
// a 4-wide register window along the kernel row, so each LDS read feeds four
// FMAs.  The slice's taps sit in LDS beside the tile (broadcast reads).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "sid_internal.h"

namespace sid {

constexpr int kCvTx = 64, kCvTy = 16;

__global__ __launch_bounds__(256) void k_convn_same(const float *__restrict__ V, int X, int Y, int Z,
                                                    const float *__restrict__ k, int kx, int ky, int kz,
                                                    float *__restrict__ out)
{
    extern __shared__ float tile[];
    const int WX = kCvTx + kx - 1, WY = kCvTy + ky - 1;
    float *ks = tile + WX * WY;   // taps of the current slice, [j][i]
    const int x0 = blockIdx.x * kCvTx, y0 = blockIdx.y * kCvTy, z = blockIdx.z;
    const int ox = kx / 2, oy = ky / 2, oz = kz / 2;
    const int tx = (threadIdx.x & 15) * 4, ty = threadIdx.x >> 4;   // 16 x 16 threads, 4 x outputs each
    // tile element (u, w) holds V[x0 + ox - (kx - 1) + u, y0 + oy - (ky - 1) + w]; output (tx + m, ty)
    // with tap (i, j) reads tile (tx + m + kx - 1 - i, ty + ky - 1 - j)
    const int bx = x0 + ox - (kx - 1), by = y0 + oy - (ky - 1);
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int l = 0; l < kz; ++l) {
        const int zi = z + oz - l;
        if (zi < 0 || zi >= Z) continue;   // block-uniform
        const float *src = V + (int64_t)zi * X * Y;
        __syncthreads();
        for (int e = threadIdx.x; e < WX * WY; e += 256) {
            const int w = e / WX, u = e - w * WX;
            const int xs = bx + u, ys = by + w;
            tile[e] = (xs >= 0 && xs < X && ys >= 0 && ys < Y) ? src[(int64_t)ys * X + xs] : 0.0f;
        }
        for (int e = threadIdx.x; e < kx * ky; e += 256) ks[e] = k[(int64_t)l * kx * ky + e];
        __syncthreads();
        for (int j = 0; j < ky; ++j) {
            const float *row = tile + (ty + ky - 1 - j) * WX + tx + kx - 1;   // row[m - i] for output tx + m
            const float *kr = ks + j * kx;
            float w0 = row[0], w1 = row[1], w2 = row[2], w3 = row[3];
            for (int i = 0; i < kx; ++i) {
                const float kv = kr[i];
                acc[0] += w0 * kv;
                acc[1] += w1 * kv;
                acc[2] += w2 * kv;
                acc[3] += w3 * kv;
                // slide: next i reads row[m - i - 1]
                w3 = w2;
                w2 = w1;
                w1 = w0;
                w0 = i + 1 < kx ? row[-(i + 1)] : 0.0f;
            }
        }
    }
    const int y = y0 + ty;
    if (y >= Y) return;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int x = x0 + tx + m;
        if (x < X) out[((int64_t)z * Y + y) * X + x] = acc[m];
    }
}

int convn_same(const float *in, const int64_t vdims[3], const float *k_dev, const int kd[3], float *out,
               hipStream_t st)
{
    if (vdims[0] < 1 || vdims[1] < 1 || vdims[2] < 1) {
        set_error("convn: volume dimensions must be positive");
        return SID_ERR_DIMS;
    }
    if (kd[0] < 1 || kd[1] < 1 || kd[2] < 1 || kd[0] > kConvnMaxXY || kd[1] > kConvnMaxXY || kd[2] > kConvnMaxZ) {
        set_error("convn: kernel %d x %d x %d outside 1..%d x 1..%d x 1..%d", kd[0], kd[1], kd[2], kConvnMaxXY,
                  kConvnMaxXY, kConvnMaxZ);
        return SID_ERR_DIMS;
    }
    if (vdims[0] > (1 << 24) || vdims[1] > (1 << 24) || vdims[2] > 65535) {
        set_error("convn: volume too large");
        return SID_ERR_DIMS;
    }
    const size_t lds = sizeof(float) * ((size_t)(kCvTx + kd[0] - 1) * (kCvTy + kd[1] - 1) + (size_t)kd[0] * kd[1]);
    const dim3 grid((unsigned)((vdims[0] + kCvTx - 1) / kCvTx), (unsigned)((vdims[1] + kCvTy - 1) / kCvTy),
                    (unsigned)vdims[2]);
    hipLaunchKernelGGL(k_convn_same, grid, dim3(256), lds, st, in, (int)vdims[0], (int)vdims[1], (int)vdims[2], k_dev,
                       kd[0], kd[1], kd[2], out);
    SID_TRY_HIP(hipGetLastError());
    return SID_OK;
}

}  // namespace sid
