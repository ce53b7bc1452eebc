// Library of normalised neuron footprints: src/generate_LFM_library_CPU.m:18-70.
//
// Per centre: the ball (or a given kernel) is voxelised as the reference does
// (meshgrid / find / ind2sub with MATLAB colon and round semantics,
// :29-50), the voxels inside the volume are painted into their bounding box,
// every box is projected in one batched launch (library_forward: the same
// persistent tap-major GEMM as the GPU library), and each patch is divided by
// its FP64 L2 norm (:51-53).  The result is the sparse matrix of :59-70 in
// compressed-row form, columns ascending as find() returns them.
//
// The patch of project_forward_patched (conv2 'full' pasted at the patch
// centre, src/project_forward_patched.m:37-73) equals the 'same'-aligned
// sensor projection on the pasted rectangle when Nnum and the PSF side are
// odd; other parities shift it by one pixel and are rejected.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "sid_internal.h"

namespace sid {

// Number of elements of MATLAB a:b (unit step): fix(b - a) + 1, with a
// difference within a few ulps below an integer counted as that integer.
static int64_t colon_count(double a, double b)
{
    const double q = b - a;
    double m = std::floor(q);
    if (q - m > 1.0 - 4.0 * 2.220446049250313e-16 * std::max(1.0, std::fabs(q))) m += 1.0;
    return m < 0 ? 0 : (int64_t)m + 1;
}

// The voxels of one footprint shape relative to its box: BWW rows (1-based i,
// j, k of ind2sub([2r+1, 2r+1, 2rr+1], find(W))) and the values W(BWW).
struct NsfShape {
    std::vector<double> bww;   // 3 per voxel
    std::vector<float> w;
    double r = 0, rr = 0;
};

static int nsf_shape(double r, const double dim[3], const float *kernel, const int64_t kd[3], NsfShape *sh)
{
    sh->bww.clear();
    sh->w.clear();
    if (kernel) {
        // generate_LFM_library_CPU.m:29-32: W = kernel, r = (size(W,1)-1)/2, rr = (size(W,3)-1)/2
        if (kd[0] < 1 || kd[1] != kd[0] || kd[2] < 1) {
            set_error("the neuron kernel must be [k, k, kz] with k >= 1 (ind2sub takes size(W,1) for both lateral axes)");
            return SID_ERR_DIMS;
        }
        sh->r = (double)(kd[0] - 1) / 2.0;
        sh->rr = (double)(kd[2] - 1) / 2.0;
        const int64_t s1 = kd[0], s12 = kd[0] * kd[0];
        const int64_t nel = kd[0] * kd[1] * kd[2];
        for (int64_t e = 0; e < nel; ++e) {
            if (kernel[e] == 0.0f) continue;    // find(W): nonzeros (NaN included), column-major
            const int64_t k = e / s12, rem = e - k * s12, j = rem / s1, i = rem - j * s1;
            sh->bww.push_back((double)(i + 1));
            sh->bww.push_back((double)(j + 1));
            sh->bww.push_back((double)(k + 1));
            sh->w.push_back(kernel[e]);
        }
        return SID_OK;
    }
    // :18,34-35: [X,Y,Z] = meshgrid(-r:r,-r:r,-rr:rr); W = (X.^2 + Y.^2 + (Z*r/rr).^2) <= r^2
    if (!(r >= 0) || r != std::floor(r) || !(dim && dim[2] > 0)) {
        set_error("the ball radius must be a whole number >= 0 and dim(3) > 0 (sid_main.m passes round(neur_rad))");
        return SID_ERR_ARG;
    }
    const double rr = r / dim[2];
    sh->r = r;
    sh->rr = rr;
    const int64_t nl = colon_count(-r, r), nz = colon_count(-rr, rr);
    for (int64_t k = 0; k < nz; ++k) {
        const double Z = -rr + (double)k;
        const double zs = Z * r / rr;
        for (int64_t j = 0; j < nl; ++j) {            // meshgrid: the second index runs over X
            const double Xv = -r + (double)j;
            for (int64_t i = 0; i < nl; ++i) {        // the first over Y
                const double Yv = -r + (double)i;
                if (Xv * Xv + Yv * Yv + zs * zs <= r * r) {
                    // ind2sub with siz(1) = 2r+1 = nl: the subscripts are the grid indices
                    sh->bww.push_back((double)(i + 1));
                    sh->bww.push_back((double)(j + 1));
                    sh->bww.push_back((double)(k + 1));
                    sh->w.push_back(1.0f);
                }
            }
        }
    }
    return SID_OK;
}

// One centre's voxels inside the volume (:43-50): bbb = round(BWW - [(r+1) (r+1) (rr+1)] + centre)
// (round half away from zero), kept when 0 < bbb <= (m1, m2, Nz); box = their bounds.
struct NsfSeed {
    std::vector<int64_t> vox;   // 3 per voxel, 1-based
    std::vector<float> val;
    int64_t box[6] = {0, 0, 0, 0, 0, 0};
};

static void nsf_place(const NsfShape &sh, const double c[3], const int64_t m[2], int64_t nzH, NsfSeed *sd)
{
    sd->vox.clear();
    sd->val.clear();
    const double shift[3] = {sh.r + 1.0, sh.r + 1.0, sh.rr + 1.0};
    const int64_t lim[3] = {m[0], m[1], nzH};
    const size_t nv = sh.w.size();
    for (size_t v = 0; v < nv; ++v) {
        int64_t b[3];
        bool in = true;
        for (int d = 0; d < 3; ++d) {
            b[d] = (int64_t)std::round((sh.bww[3 * v + d] - shift[d]) + c[d]);
            in = in && b[d] > 0 && b[d] <= lim[d];
        }
        if (!in) continue;
        for (int d = 0; d < 3; ++d) sd->vox.push_back(b[d]);
        sd->val.push_back(sh.w[v]);
    }
    if (sd->val.empty()) return;
    for (int d = 0; d < 3; ++d) {
        sd->box[2 * d] = sd->box[2 * d + 1] = sd->vox[d];
        for (size_t v = 1; v < sd->val.size(); ++v) {
            sd->box[2 * d] = std::min(sd->box[2 * d], sd->vox[3 * v + d]);
            sd->box[2 * d + 1] = std::max(sd->box[2 * d + 1], sd->vox[3 * v + d]);
        }
    }
}

static int nsf_prepare(const sid_psf *p, int64_t ns, const double *centers, double r, const double dim[3],
                       const int64_t m[2], const float *kernel, const int64_t kd[3], std::vector<NsfSeed> *seeds)
{
    if (p->cropped) {
        set_error("the footprint library projects with the full PSF (psf_ballistic.H); create the handle without CAindex");
        return SID_ERR_STATE;
    }
    if (p->n % 2 == 0 || p->M % 2 == 0) {
        set_error("the footprint library needs odd Nnum and PSF side (project_forward_patched's paste is shifted by "
                  "one pixel otherwise)");
        return SID_ERR_DIMS;
    }
    if (m[0] < 1 || m[1] < 1) {
        set_error("volume size must be positive");
        return SID_ERR_DIMS;
    }
    NsfShape sh;
    SID_TRY(nsf_shape(r, dim, kernel, kd, &sh));
    seeds->assign(ns, NsfSeed());
    for (int64_t s = 0; s < ns; ++s) {
        const double c[3] = {centers[s], centers[s + ns], centers[s + 2 * ns]};   // MATLAB centers(s, :)
        nsf_place(sh, c, m, p->nzH, &(*seeds)[s]);
    }
    return SID_OK;
}

}  // namespace sid

using namespace sid;

extern "C" {

int sid_library_nsf_bound(const sid_psf *p, int64_t ns, const double *centers, double r, const double dim[3],
                          const int64_t m[2], const float *kernel, const int64_t kdims[3], int64_t *bound)
{
    if (!p || ns < 0 || (ns > 0 && !centers) || !m || !bound || (kernel && !kdims)) {
        set_error("sid_library_nsf_bound: null argument");
        return SID_ERR_ARG;
    }
    std::vector<NsfSeed> seeds;
    SID_TRY(nsf_prepare(p, ns, centers, r, dim, m, kernel, kdims, &seeds));
    int64_t tot = 0;
    for (const NsfSeed &sd : seeds) {
        if (sd.val.empty()) continue;
        int64_t rect[4];
        patched_rect(p->n, p->M, m[0], m[1], sd.box, rect);
        tot += (rect[1] - rect[0] + 1) * (rect[3] - rect[2] + 1);
    }
    *bound = tot;
    return SID_OK;
}

int sid_library_nsf(sid_psf *p, int64_t ns, const double *centers, double r, const double dim[3], const int64_t m[2],
                    const float *kernel, const int64_t kdims[3], int64_t *row_ptr, int64_t *cols, double *vals,
                    int64_t cap)
{
    if (!p || ns < 0 || (ns > 0 && !centers) || !m || !row_ptr || (kernel && !kdims) || cap < 0 ||
        (cap > 0 && (!cols || !vals))) {
        set_error("sid_library_nsf: null argument");
        return SID_ERR_ARG;
    }
    if (p->kern) {
        set_error("the footprint library projects with the PSF alone; remove the handle's neuron kernel");
        return SID_ERR_ARG;
    }
    std::vector<NsfSeed> seeds;
    SID_TRY(nsf_prepare(p, ns, centers, r, dim, m, kernel, kdims, &seeds));
    // seeds with voxels inside the volume, their boxes and painted values
    std::vector<int64_t> live, boxes, voff(1, 0), rect(4 * ns, 0), poff(1, 0);
    std::vector<float> hv;
    for (int64_t s = 0; s < ns; ++s) {
        const NsfSeed &sd = seeds[s];
        if (sd.val.empty()) continue;
        live.push_back(s);
        const int64_t *b = sd.box;
        const int64_t wx = b[1] - b[0] + 1, wy = b[3] - b[2] + 1, wz = b[5] - b[4] + 1;
        const size_t o = hv.size();
        hv.resize(o + wx * wy * wz, 0.0f);
        for (size_t v = 0; v < sd.val.size(); ++v) {   // later voxels overwrite, as the paint loop of :28-30
            const int64_t *q = &sd.vox[3 * v];
            hv[o + (q[0] - b[0]) + wx * ((q[1] - b[2]) + wy * (q[2] - b[4]))] = sd.val[v];
        }
        voff.push_back((int64_t)hv.size());
        boxes.insert(boxes.end(), b, b + 6);
        patched_rect(p->n, p->M, m[0], m[1], b, &rect[4 * s]);
        const int64_t *rc = &rect[4 * s];
        poff.push_back(poff.back() + (rc[1] - rc[0] + 1) * (rc[3] - rc[2] + 1));
    }
    const int64_t nl = (int64_t)live.size();
    if (poff.back() > cap) {
        set_error("sid_library_nsf: %lld entries may be needed, capacity %lld (sid_library_nsf_bound)",
                  (long long)poff.back(), (long long)cap);
        return SID_ERR_ARG;
    }
    std::vector<float> patches(poff.back());
    if (nl > 0) {
        SID_TRY_HIP(hipSetDevice(p->device));
        struct Buf {
            void *ptr = nullptr;
            ~Buf()
            {
                if (ptr) (void)hipFree(ptr);
            }
        } dv, dout;
        SID_TRY_HIP(hipMalloc(&dv.ptr, sizeof(float) * std::max<size_t>(hv.size(), 1)));
        SID_TRY_HIP(hipMalloc(&dout.ptr, sizeof(float) * std::max<int64_t>(poff.back(), 1)));
        SID_TRY(copy_h2d(dv.ptr, hv.data(), sizeof(float) * hv.size(), p->stream));
        SID_TRY(library_forward(p, nl, boxes.data(), (const float *)dv.ptr, voff.data(), m[0], m[1],
                                (float *)dout.ptr, p->stream, 0));
        SID_TRY(copy_d2h(patches.data(), dout.ptr, sizeof(float) * poff.back(), p->stream));
        lib_plan_free(p);   // the cached plan keys on these buffers, freed on return
    }
    // Q = Q / norm(Q) (:52-53) and the sparse row of find(Q) (:54-55, :63-70)
    int64_t nnz = 0, li = 0;
    row_ptr[0] = 0;
    for (int64_t s = 0; s < ns; ++s) {
        if (!seeds[s].val.empty()) {
            const float *pt = patches.data() + poff[li];
            const int64_t np = poff[li + 1] - poff[li];
            ++li;
            double ss = 0.0;
            for (int64_t e = 0; e < np; ++e) ss += (double)pt[e] * (double)pt[e];
            const double nrm = std::sqrt(ss);
            const int64_t *rc = &rect[4 * s];
            const int64_t w = rc[1] - rc[0] + 1;
            if (nrm > 0.0) {
                for (int64_t e = 0; e < np; ++e) {
                    if (pt[e] == 0.0f) continue;
                    const int64_t x = rc[0] - 1 + e % w, y = rc[2] - 1 + e / w;
                    cols[nnz] = x + m[0] * y;
                    vals[nnz] = (double)pt[e] / nrm;
                    ++nnz;
                }
            }
        }
        row_ptr[s + 1] = nnz;
    }
    return SID_OK;
}

}  // extern "C"
