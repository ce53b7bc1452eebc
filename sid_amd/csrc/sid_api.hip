// C ABI of libsidproj: handle management, the MEX-shaped projector entry
// points and the on-device ISRA loop (external/LFrecon_util/deconvRL.m).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <mutex>
#include <set>
#include <vector>

#include "sid_internal.h"

namespace sid {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

static int supported_np(int64_t nn2)
{
    static const int nts[] = {1, 2, 4, 8, 11, 15, 16};
    const int64_t need = (nn2 + 15) / 16;
    for (int nt : nts)
        if (nt >= need) return nt * 16;
    return 0;
}

static bool is_gfx950(int dev)
{
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return false;
    return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
}

// Event pool used for the per-phase timing of the last call.
struct Timer {
    sid_psf *p;
    struct Span { int phase, e0, e1; };
    std::vector<Span> spans;
    std::vector<hipEvent_t> evs;
    hipStream_t st;
    explicit Timer(sid_psf *p_, hipStream_t s) : p(p_), st(s) {}
    ~Timer()
    {
        for (auto e : evs) (void)hipEventDestroy(e);
    }
    // event recorded on s (default: the timer's stream); -1 when no event could be made
    int mark(hipStream_t s = nullptr)
    {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        evs.push_back(e);
        (void)hipEventRecord(e, s ? s : st);
        return (int)evs.size() - 1;
    }
    void span(int phase, int e0) { spans.push_back({phase, e0, e0 + 1}); }
    void span(int phase, int e0, int e1) { spans.push_back({phase, e0, e1}); }
    void finish(int e_start, int e_end)
    {
        for (int i = 0; i < 12; ++i) p->timing[i] = 0.0;
        fft_timing_collect(p, &p->timing[7], &p->timing[8]);
        for (auto &s : spans) {
            float ms = 0.f;
            if (s.e0 >= 0 && s.e1 >= 0 && hipEventElapsedTime(&ms, evs[s.e0], evs[s.e1]) == hipSuccess)
                p->timing[s.phase] += ms;
        }
        float ms = 0.f;
        if (e_start >= 0 && e_end >= 0 && hipEventElapsedTime(&ms, evs[e_start], evs[e_end]) == hipSuccess)
            p->timing[4] = ms;
    }
};

static int create_common(sid_psf *p, const int64_t dims[5], const int32_t *CA, bool full)
{
    const int64_t M = dims[0], M2 = dims[1], n = dims[2], n2 = dims[3], nz = dims[4];
    if (M < 1 || M2 != M || n < 1 || n2 != n || nz < 1) {
        set_error("H must be [M, M, Nnum, Nnum, Nz] with square kernels and square lenslets (got %lld x %lld x %lld x %lld x %lld)",
                  (long long)M, (long long)M2, (long long)n, (long long)n2, (long long)nz);
        return SID_ERR_DIMS;
    }
    if (nz > 512) {
        set_error("at most 512 planes are supported (got %lld)", (long long)nz);
        return SID_ERR_DIMS;
    }
    const int np = supported_np(n * n);
    if (np == 0) {
        set_error("Nnum = %lld exceeds the supported lenslet size (<= 16)", (long long)n);
        return SID_ERR_DIMS;
    }
    p->M = M; p->n = n; p->nzH = nz;
    p->S = (n * n + 7) / 8 * 8;
    p->NP = np;
    p->planes.assign(nz, Plane{});
    if (CA) {
        for (int64_t c = 0; c < nz; ++c) {
            const int lo = CA[c] - 1, hi = CA[c + nz] - 1;
            if (lo < 0 || hi >= M || hi < lo) {
                set_error("CAindex(%lld,:) = [%d %d] is not a crop inside 1..%lld", (long long)c + 1, lo + 1, hi + 1,
                          (long long)M);
                return SID_ERR_DIMS;
            }
        }
        p->cropped = true;
    }
    p->full_support = full;
    if (full) {
        for (auto &pl : p->planes) { pl.lo = 0; pl.hi = (int)M - 1; pl.empty = false; }
    } else {
        SID_TRY(support_scan(p));
    }
    for (int64_t c = 0; c < nz; ++c) {
        Plane &pl = p->planes[c];
        if (CA) {
            const int clo = CA[c] - 1, chi = CA[c + nz] - 1, Mc = chi - clo + 1;
            pl.off_f = clo + Mc / 2;            // conv2 'same' offset of the crop (forwardProjectACC.m:25)
            pl.off_b = clo + (Mc - 1 - Mc / 2);  // after imrotate(., 180) (backwardProjectACC_new.m:21)
            if (!pl.empty) {
                pl.lo = std::max(pl.lo, clo);
                pl.hi = std::min(pl.hi, chi);
                if (pl.hi < pl.lo) pl.empty = true;
            }
        } else {
            pl.off_f = (int)(M / 2);                 // conv2FFT_.m:6,10-12
            pl.off_b = (int)(M - 1 - M / 2);         // rot180 kernel, backwardProjectGPU_new.m:19
        }
    }
    return build_weights(p, 0);
}

static void destroy(sid_psf *p)
{
    if (!p) return;
    (void)hipSetDevice(p->device);
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    SensorPlan &sp = p->plan;
    lib_plan_free(p);
    fft_state_free(sp.fft);
    free_plan_buffers(sp);
    for (void *q : {(void *)sp.fwd.d_segs, (void *)sp.fwd.d_items, (void *)sp.bwd.d_segs, (void *)sp.bwd.d_items,
                    (void *)sp.fwd.d_tmap, (void *)sp.bwd.d_tmap, (void *)sp.fwd.d_steps, (void *)sp.bwd.d_steps,
                    (void *)sp.fwd.d_run0, (void *)sp.bwd.d_run0, (void *)sp.bwd.d_plane_slab0, (void *)sp.fwd.d_plane_slab0, (void *)p->H, (void *)p->wf,
                    (void *)p->wb, (void *)p->wf_hi, (void *)p->wb_hi, (void *)p->wf_lo, (void *)p->wb_lo,
                    (void *)p->wf_d, (void *)p->wb_d, (void *)p->kern, (void *)p->kwork, (void *)p->kfrag,
                    (void *)p->hbuf[0], (void *)p->hbuf[1], (void *)p->nf})
        if (q) (void)hipFree(q);
    if (p->kfft) {
        convn_fft_free(*p->kfft);
        delete p->kfft;
    }
    if (p->aux) {
        (void)hipStreamSynchronize(p->aux);
        (void)hipStreamDestroy(p->aux);
    }
    if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
    if (p->ev_null) (void)hipEventDestroy(p->ev_null);
    if (p->ev_join) (void)hipEventDestroy(p->ev_join);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

static int check_device()
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_error("no HIP device visible");
        return SID_ERR_NODEV;
    }
    int dev = 0;
    SID_TRY_HIP(hipGetDevice(&dev));
    if (!is_gfx950(dev)) {
        set_error("device %d is not gfx950 (MI355X); this library is built for gfx950 only", dev);
        return SID_ERR_NODEV;
    }
    return SID_OK;
}

// The stream a device-pointer entry point runs on.  A caller's stream is used
// as given.  NULL is the device's default (null) stream, as in the HIP runtime:
// the work then runs on the handle's own stream (non-blocking, so it would not
// wait for the default stream by itself) after an event that makes it wait for
// everything already queued on the default stream (the caller's producers of
// the inputs).  Every entry point is synchronous, so its outputs are complete
// when it returns either way.
static int pick_stream(sid_psf *p, void *stream, hipStream_t *st)
{
    if (stream) {
        *st = (hipStream_t)stream;
        return SID_OK;
    }
    if (!p->ev_null) SID_TRY_HIP(hipEventCreateWithFlags(&p->ev_null, hipEventDisableTiming));
    SID_TRY_HIP(hipEventRecord(p->ev_null, nullptr));
    SID_TRY_HIP(hipStreamWaitEvent(p->stream, p->ev_null, 0));
    *st = p->stream;
    return SID_OK;
}

}  // namespace sid

using namespace sid;

extern "C" {

const char *sid_last_error(void) { return g_err; }
int sid_version(void) { return 1; }

int sid_device_count(void)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess) return 0;
    int k = 0;
    for (int d = 0; d < ndev; ++d) k += is_gfx950(d) ? 1 : 0;
    return k;
}

// Handles alive in this process (for callers that pass them through an integer,
// such as a MATLAB uint64, and need to check one before use).
static std::mutex g_live_mu;
static std::set<const sid_psf *> g_live;

static int psf_create_impl(const float *H, bool host, const int64_t dims[5], const int32_t *CA, bool full,
                           sid_psf **out)
{
    if (!H || !dims || !out) {
        set_error("sid_psf_create: null argument");
        return SID_ERR_ARG;
    }
    *out = nullptr;
    SID_TRY(check_device());
    sid_psf *p = new (std::nothrow) sid_psf();
    if (!p) {
        set_error("host allocation failed");
        return SID_ERR_OOM;
    }
    int rc = SID_OK;
    do {
        if (hipGetDevice(&p->device) != hipSuccess) { rc = SID_ERR_HIP; set_error("hipGetDevice failed"); break; }
        if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
            rc = SID_ERR_HIP;
            set_error("hipStreamCreate failed");
            break;
        }
        const char *ser = std::getenv("SID_NO_OVERLAP");
        p->overlap = !(ser && ser[0] && ser[0] != '0');
        const char *gb = std::getenv("SID_GEMM_BLOCKS");
        p->gemm_blocks = gb ? std::max(0, std::atoi(gb)) : 0;
        p->gemm_blocks_env = p->gemm_blocks > 0;
        if (p->gemm_blocks == 0) {
            hipDeviceProp_t prop;
            p->gemm_blocks = hipGetDeviceProperties(&prop, p->device) == hipSuccess ? prop.multiProcessorCount : 256;
        }
        for (int i = 0; i < 5; ++i)
            if (dims[i] < 1) { rc = SID_ERR_DIMS; set_error("H has an empty dimension"); }
        if (rc) break;
        if (hipMalloc(&p->nf, 2 * sizeof(unsigned long long)) != hipSuccess ||
            hipMemset(p->nf, 0, 2 * sizeof(unsigned long long)) != hipSuccess) {
            rc = SID_ERR_OOM;
            set_error("hipMalloc of the non-finite flags failed");
            break;
        }
        const int64_t elems = dims[0] * dims[1] * dims[2] * dims[3] * dims[4];
        hipError_t e = hipMalloc(&p->H, elems * sizeof(float));
        if (e != hipSuccess) { rc = SID_ERR_OOM; set_error("hipMalloc(H, %lld B) failed", (long long)(elems * 4)); break; }
        if (host) {
            rc = copy_h2d(p->H, H, elems * sizeof(float), nullptr);
            if (rc == SID_OK && (e = hipStreamSynchronize(nullptr)) != hipSuccess) rc = SID_ERR_HIP;
            if (rc) { set_error("H upload failed"); break; }
        } else {
            e = hipMemcpy(p->H, H, elems * sizeof(float), hipMemcpyDeviceToDevice);
            if (e != hipSuccess) { rc = SID_ERR_HIP; set_error("H upload failed: %s", hipGetErrorString(e)); break; }
        }
        rc = create_common(p, dims, CA, full);
    } while (0);
    if (rc != SID_OK) {
        destroy(p);
        return rc;
    }
    {
        std::lock_guard<std::mutex> lk(g_live_mu);
        g_live.insert(p);
    }
    *out = p;
    return SID_OK;
}

int sid_psf_create(const float *H, const int64_t dims[5], const int32_t *CAindex, sid_psf **out)
{
    return psf_create_impl(H, true, dims, CAindex, false, out);
}

int sid_psf_create_device(const float *H_dev, const int64_t dims[5], const int32_t *CAindex, sid_psf **out)
{
    return psf_create_impl(H_dev, false, dims, CAindex, false, out);
}

int sid_psf_create_ex(const float *H, const int64_t dims[5], const int32_t *CAindex, int flags, sid_psf **out)
{
    return psf_create_impl(H, (flags & SID_FLAG_H_ON_DEVICE) == 0, dims, CAindex, (flags & SID_FLAG_FULL_SUPPORT) != 0,
                           out);
}

int sid_psf_update(sid_psf *p, const float *H, int on_device)
{
    if (!p || !H) { set_error("sid_psf_update: null argument"); return SID_ERR_ARG; }
    SID_TRY_HIP(hipSetDevice(p->device));
    const int64_t elems = p->M * p->M * p->n * p->n * p->nzH;
    if (on_device)
        SID_TRY_HIP(hipMemcpyAsync(p->H, H, elems * sizeof(float), hipMemcpyDeviceToDevice, p->stream));
    else
        SID_TRY(copy_h2d(p->H, H, elems * sizeof(float), p->stream));
    SID_TRY_HIP(hipStreamSynchronize(p->stream));
    const bool had[4] = {p->wf != nullptr, p->wf_hi != nullptr, p->wf_lo != nullptr, p->wf_d != nullptr};
    lib_plan_free(p);  // its captured graph holds the old weight pointers
    free_weights(p);
    for (int w = 0; w < 4; ++w)
        if (had[w] || w == 0) SID_TRY(build_weights(p, w));
    // the sensor plan holds values derived from H (the FFT planes' spectra, the
    // GEMM's zero-tile tables): the next projection builds it again
    fft_state_free(p->plan.fft);
    free_plan_buffers(p->plan);
    return SID_OK;
}

int sid_poisson_nll_grad_device(sid_psf *p, const float *V, const float *y, const int64_t dims[4], double *nll,
                                float *gV, float *gH, void *stream)
{
    if (!p || !V || !y || !dims) { set_error("sid_poisson_nll_grad_device: null argument"); return SID_ERR_ARG; }
    if (dims[0] < 1 || dims[1] < 1 || dims[2] < 1 || dims[3] < 1) {
        set_error("dimensions must be positive");
        return SID_ERR_DIMS;
    }
    if (dims[2] != p->nzH) {
        set_error("the likelihood needs volumes with size(H,5) = %lld planes", (long long)p->nzH);
        return SID_ERR_NZ;
    }
    if (p->kern) { set_error("gradients are not available on a handle with a neuron kernel"); return SID_ERR_ARG; }
    SID_TRY_HIP(hipSetDevice(p->device));
    hipStream_t st;
    SID_TRY(pick_stream(p, stream, &st));
    return grad_core(p, V, y, dims, 0, nll, gV, gH, st);
}

int sid_backward_ones_device(sid_psf *p, int64_t X, int64_t Y, double *B1, void *stream)
{
    if (!p || !B1) { set_error("sid_backward_ones_device: null argument"); return SID_ERR_ARG; }
    if (X < 1 || Y < 1 || X > INT32_MAX || Y > INT32_MAX) { set_error("sensor size out of range"); return SID_ERR_DIMS; }
    if (p->kern) { set_error("B(1) is taken of the projector alone; remove the neuron kernel"); return SID_ERR_ARG; }
    SID_TRY_HIP(hipSetDevice(p->device));
    hipStream_t st;
    SID_TRY(pick_stream(p, stream, &st));
    return backward_ones_f64(p, X, Y, B1, st);
}

int sid_project_vjp_device(sid_psf *p, const float *V, const float *G, const int64_t dims[4], float *gV, float *gH,
                           void *stream)
{
    if (!p || !G || !dims || (gH && !V)) { set_error("sid_project_vjp_device: null argument"); return SID_ERR_ARG; }
    if (dims[0] < 1 || dims[1] < 1 || dims[2] < 1 || dims[3] < 1) {
        set_error("dimensions must be positive");
        return SID_ERR_DIMS;
    }
    if (dims[2] != p->nzH) {
        set_error("the projector gradient needs volumes with size(H,5) = %lld planes", (long long)p->nzH);
        return SID_ERR_NZ;
    }
    if (p->kern) { set_error("gradients are not available on a handle with a neuron kernel"); return SID_ERR_ARG; }
    SID_TRY_HIP(hipSetDevice(p->device));
    hipStream_t st;
    SID_TRY(pick_stream(p, stream, &st));
    return grad_core(p, V, G, dims, 1, nullptr, gV, gH, st);
}

int sid_psf_destroy(sid_psf *psf)
{
    if (psf) {
        std::lock_guard<std::mutex> lk(g_live_mu);
        g_live.erase(psf);
    }
    destroy(psf);
    return SID_OK;
}

int sid_psf_is_live(const sid_psf *psf)
{
    std::lock_guard<std::mutex> lk(g_live_mu);
    if (psf && g_live.count(psf)) return SID_OK;
    set_error("not a live sid_psf handle");
    return SID_ERR_ARG;
}

int sid_psf_set_precision(sid_psf *p, int precision)
{
    if (!p) { set_error("null handle"); return SID_ERR_ARG; }
    if (precision < SID_PREC_FP32 || precision > SID_PREC_FP64) {
        set_error("unknown precision %d", precision);
        return SID_ERR_ARG;
    }
    SID_TRY_HIP(hipSetDevice(p->device));
    p->precision = precision;
    if (precision == SID_PREC_BF16X2 || precision == SID_PREC_BF16X3) SID_TRY(build_weights(p, 1));
    if (precision == SID_PREC_BF16X3) SID_TRY(build_weights(p, 2));
    if (precision == SID_PREC_FP64) SID_TRY(build_weights(p, 3));
    if (p->plan.vol) SID_TRY(ensure_plan(p, p->plan.X, p->plan.Y, p->plan.nz));
    return SID_OK;
}

int sid_psf_set_method(sid_psf *p, int method)
{
    if (!p) { set_error("null handle"); return SID_ERR_ARG; }
    if (method != SID_METHOD_AUTO && method != SID_METHOD_DIRECT && method != SID_METHOD_FFT) {
        set_error("unknown method %d", method);
        return SID_ERR_ARG;
    }
    p->method = method;
    return SID_OK;
}

int sid_psf_set_overlap(sid_psf *p, int on)
{
    if (!p) { set_error("null handle"); return SID_ERR_ARG; }
    const bool v = on != 0;
    if (v == p->overlap) return SID_OK;
    if (p->plan.vol) SID_TRY_HIP(hipStreamSynchronize(p->stream));
    if (p->aux) SID_TRY_HIP(hipStreamSynchronize(p->aux));
    p->overlap = v;
    fft_state_free(p->plan.fft);   // the sensor plan is built again by the next projection
    free_plan_buffers(p->plan);
    return SID_OK;
}

int sid_psf_set_plane_methods(sid_psf *p, const int32_t *methods, int64_t n)
{
    if (!p) { set_error("null handle"); return SID_ERR_ARG; }
    if (!methods) {
        p->pin.clear();
        return SID_OK;
    }
    if (n != p->nzH) {
        set_error("sid_psf_set_plane_methods: %lld entries for %lld planes", (long long)n, (long long)p->nzH);
        return SID_ERR_ARG;
    }
    std::vector<char> pin(n, 0);
    for (int64_t c = 0; c < n; ++c) {
        const int m = methods[c];
        if (m != SID_METHOD_AUTO && m != SID_METHOD_DIRECT && m != SID_METHOD_FFT) {
            set_error("plane %lld: unknown method %d", (long long)c + 1, m);
            return SID_ERR_ARG;
        }
        if (m == SID_METHOD_FFT && !fft_eligible(p, c)) {
            set_error("plane %lld cannot take the FFT path (empty, or an even support whose adjoint is not the "
                      "reference backward)", (long long)c + 1);
            return SID_ERR_ARG;
        }
        pin[c] = (char)m;
    }
    p->pin = pin;
    return SID_OK;
}

int sid_psf_get_info(const sid_psf *p, sid_psf_info *info)
{
    if (!p || !info) { set_error("null argument"); return SID_ERR_ARG; }
    std::memset(info, 0, sizeof(*info));
    info->M = p->M; info->Nnum = p->n; info->Nz = p->nzH; info->taps_max = p->T;
    int64_t wb = (p->wf_elems + p->wb_elems) * 4;
    if (p->wf_hi) wb += (p->wf2_elems + p->wb2_elems) * 2;
    if (p->wf_lo) wb += (p->wf2_elems + p->wb2_elems) * 2;
    if (p->wf_d) wb += (p->wf_elems + p->wb_elems) * 8;
    info->weight_bytes = wb;
    for (int64_t c = 0; c < p->nzH && c < 512; ++c) {
        info->support_lo[c] = p->planes[c].empty ? 0 : p->planes[c].lo;
        info->support_hi[c] = p->planes[c].empty ? -1 : p->planes[c].hi;
        const bool fft = (size_t)c < p->plan.fft_use.size() && p->plan.fft_use[c];
        info->plane_method[c] = p->planes[c].empty ? 0 : (fft ? SID_METHOD_FFT : SID_METHOD_DIRECT);
    }
    for (int i = 0; i < 4; ++i) info->autotune_ms[i] = p->auto_ms[i];
    info->spectra_bytes = p->plan.fft.g_bytes;
    {
        const FftState &f = p->plan.fft;
        const bool t = f.active && f.tiled && !f.tg.empty();
        const TfGroup *w = t ? &f.tg.back() : nullptr;   // the widest group
        info->fft_tile[0] = t ? w->T : 0;
        info->fft_tile[1] = t ? w->B : 0;
        info->fft_tile[2] = t ? w->ntile : 0;
        info->fft_tile[3] = t ? w->K : 0;
        info->fft_groups = t ? (int32_t)f.tg.size() : 0;
        for (int gi = 0; gi < 8; ++gi) {
            const bool h = t && gi < (int)f.tg.size();
            info->fft_group_planes[gi] = h ? f.tg[gi].ns : 0;
            info->fft_group_T[gi] = h ? f.tg[gi].T : 0;
        }
        if (t) tf_gemm_work(f, info->fft_gemm_bytes, info->fft_gemm_flops);
    }
    return SID_OK;
}

static int check_vdims(sid_psf *p, const int64_t vdims[3], bool rl)
{
    if (!vdims || vdims[0] < 1 || vdims[1] < 1 || vdims[2] < 1) {
        set_error("volume dimensions must be positive");
        return SID_ERR_DIMS;
    }
    if (vdims[2] > p->nzH) {
        set_error("volume has %lld planes but H has %lld (forwardProjectGPU uses H(:,:,:,:,cc) for every plane)",
                  (long long)vdims[2], (long long)p->nzH);
        return SID_ERR_NZ;
    }
    if (rl && vdims[2] != p->nzH) {
        set_error("deconvRL volume must have size(H,5) = %lld planes (got %lld)", (long long)p->nzH,
                  (long long)vdims[2]);
        return SID_ERR_NZ;
    }
    return SID_OK;
}

// The two halves of one projection (dir 0 forward, 1 backward): the direct
// GEMM over the direct planes (MFMA-bound) and the FFT path over the others
// (bound by the spectra stream from HBM).  They read the same staged volume or
// image and write disjoint buffers (split-K slabs; FFT grid and volume spectra),
// so when both are present the GEMM runs on the handle's second stream, forked
// from st after the staging and joined back before the reduction that sums the
// two.  Phase times go to timing[ph_gemm] and timing[ph_fft].
// The backward reduction that follows a backward projection (arguments of
// launch_bwd_reduce).  With both halves present its direct planes are reduced
// on the GEMM's stream as soon as the GEMM is done, beside the rest of the FFT
// path; the caller then reduces the FFT planes only.
// When every plane is on the tiled FFT path the FFT half applies the whole
// reduction itself (*tail_all_done); the caller then launches none.
static int project_halves(sid_psf *p, int dir, hipStream_t st, Timer &tm, int ph_gemm, int ph_fft,
                          const BwdTail *tail = nullptr, bool *tail_direct_done = nullptr,
                          bool *tail_all_done = nullptr)
{
    if (tail_direct_done) *tail_direct_done = false;
    if (tail_all_done) *tail_all_done = false;
    const GemmPlan &g = dir == 0 ? p->plan.fwd : p->plan.bwd;
    if (!(p->overlap && p->plan.fft.active && !g.items.empty())) {
        const int a = tm.mark();
        SID_TRY(launch_gemm(p, dir, st));
        const int b = tm.mark();
        SID_TRY(dir == 0 ? fft_forward(p, st) : fft_backward(p, st, tail_all_done ? tail : nullptr, tail_all_done));
        const int c = tm.mark();
        tm.span(ph_gemm, a, b);
        tm.span(ph_fft, b, c);
        return SID_OK;
    }
    SID_TRY(ensure_aux(p));
    SID_TRY_HIP(hipEventRecord(p->ev_fork, st));
    SID_TRY_HIP(hipStreamWaitEvent(p->aux, p->ev_fork, 0));
    const int a = tm.mark(p->aux);
    const int rc_g = launch_gemm(p, dir, p->aux, dir == 0 && p->gemm_blocks_f > 0 ? p->gemm_blocks_f : p->gemm_blocks);
    const int b = tm.mark(p->aux);
    int rc_t = SID_OK;
    if (rc_g == SID_OK && tail && dir == 1) {
        rc_t = launch_bwd_reduce(p, tail->B_col, tail->rl_update, p->aux, tail->check, tail->mark, 1);
        if (rc_t == SID_OK && tail_direct_done) *tail_direct_done = true;
    }
    // join even after a failed launch so that st never runs ahead of work left on aux
    SID_TRY_HIP(hipEventRecord(p->ev_join, p->aux));
    const int c = tm.mark();
    const int rc_f = rc_g == SID_OK ? (dir == 0 ? fft_forward(p, st) : fft_backward(p, st)) : SID_OK;
    const int d = tm.mark();
    SID_TRY_HIP(hipStreamWaitEvent(st, p->ev_join, 0));
    SID_TRY(rc_g);
    SID_TRY(rc_t);
    SID_TRY(rc_f);
    tm.span(ph_gemm, a, b);
    tm.span(ph_fft, c, d);
    return SID_OK;
}

// Mean device time (ms) of one forward plus one backward projection, both
// halves as the projector runs them, on the current plan.  The staged volume
// and image hold whatever the plan buffers hold: no kernel's time depends on
// the values.
static double time_projection_pair(sid_psf *p, hipStream_t st, int *rc)
{
    // one warm-up pair, then the median of kRounds timed pairs (robust to a
    // slow or fast outlier on a shared device)
    constexpr int kRounds = 6;
    hipEvent_t ev[kRounds + 1] = {};
    *rc = SID_OK;
    for (auto &e : ev)
        if (*rc == SID_OK && hipEventCreate(&e) != hipSuccess) *rc = SID_ERR_HIP;
    std::vector<float> ms;
    {
        Timer tm(p, st);
        for (int r = 0; r <= kRounds && *rc == SID_OK; ++r) {
            if (r > 0 && hipEventRecord(ev[r - 1], st) != hipSuccess) *rc = SID_ERR_HIP;
            if (*rc == SID_OK) *rc = project_halves(p, 0, st, tm, 0, 5);
            if (*rc == SID_OK) *rc = project_halves(p, 1, st, tm, 2, 6);
        }
        if (*rc == SID_OK && (hipEventRecord(ev[kRounds], st) != hipSuccess ||
                              hipEventSynchronize(ev[kRounds]) != hipSuccess))
            *rc = SID_ERR_HIP;
        for (int r = 0; r < kRounds && *rc == SID_OK; ++r) {
            float t = 0.f;
            if (hipEventElapsedTime(&t, ev[r], ev[r + 1]) != hipSuccess) *rc = SID_ERR_HIP;
            ms.push_back(t);
        }
        if (*rc != SID_OK) (void)hipStreamSynchronize(st);
    }
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    if (*rc != SID_OK || ms.empty()) return 0.0;
    std::sort(ms.begin(), ms.end());
    return 0.5 * (ms[(kRounds - 1) / 2] + ms[kRounds / 2]);
}

// Split of the planes between the GEMM and the FFT path when the two halves
// overlap.  autotune_methods prices each plane alone (direct GEMM time against
// FFT-path time), which minimises the serial sum; overlapped, a projection
// costs about the longer half plus their interference, which favours giving
// the GEMM more planes.  Starting from that choice (k FFT planes) or from
// every eligible plane on the FFT path (when 1 % faster; in BF16X2 it is kept
// unless 3 % slower, and then no other split is tried), time the
// overlapped forward + backward pair with the narrowest FFT planes moved to the
// GEMM s at a time (s about a ninth of k) while that gains 1 % over the
// incumbent timed before and after it, else with s more FFT planes at a time
// on the same terms, and keep the fastest.
// Candidate plans take the FFT planes widest first.  Skipped when planes are
// pinned or the overlap is off; cached with the per-plane choice.
static int autotune(sid_psf *p, int64_t X, int64_t Y, int64_t nz, hipStream_t st)
{
    SID_TRY(autotune_methods(p, X, Y, nz, st));
    if (p->method != SID_METHOD_AUTO || p->precision == SID_PREC_FP64 || p->auto_refined) return SID_OK;
    p->auto_refined = true;
    p->gemm_blocks_f = 0;
    if (!p->overlap || !p->pin.empty() || p->auto_fft.size() != (size_t)p->nzH) return SID_OK;
    std::vector<int> order;  // eligible planes, widest support first
    for (int64_t c = 0; c < p->nzH; ++c)
        if (fft_eligible(p, c)) order.push_back((int)c);
    auto width = [&](int c) { return p->planes[c].hi - p->planes[c].lo; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return width(a) > width(b); });
    int k0 = 0;
    for (int c : order) k0 += p->auto_fft[c];
    if (k0 == 0) return SID_OK;
    const int s = std::max(1, (k0 + 4) / 9), ne = (int)order.size();
    auto use_prefix = [&](int k) {
        std::fill(p->auto_fft.begin(), p->auto_fft.end(), 0);
        for (int i = 0; i < k; ++i) p->auto_fft[order[i]] = 1;
    };
    int rc = SID_OK;
    const bool log = std::getenv("SID_AUTOTUNE_LOG") != nullptr;
    auto eval = [&](int k) {
        use_prefix(k);
        rc = ensure_plan(p, X, Y, nz);
        const double t = rc == SID_OK ? time_projection_pair(p, st, &rc) : 0.0;
        if (log) std::fprintf(stderr, "sid autotune: %d FFT planes (start %d, step %d): %.3f ms per pair\n", k, k0, s, t);
        return t;
    };
    int best_k = k0;
    double best_t = eval(k0);
    // every eligible plane on the FFT path (nothing beside it) is a candidate
    // of its own: with the tiled form, narrow planes cost little there.  In
    // the BF16X2 mode it is taken, and the search ends, unless it is 3 % slower
    // than the incumbent (timed before and after it): its planes all carry FP16
    // spectra, with a smaller error than the BF16-quantised PSF of direct
    // planes.  Otherwise it must win by 1 %, like any other move.
    bool settled = false;
    if (k0 < ne && rc == SID_OK) {
        const bool pref = p->precision == SID_PREC_BF16X2;
        const double lim = pref ? 1.03 : 0.99;
        const double t = eval(ne);
        if (rc == SID_OK && t < lim * best_t) {
            const double t_again = eval(best_k);
            if (rc == SID_OK && t < lim * t_again) {
                best_t = t;
                best_k = ne;
                settled = pref;
            }
        }
    }
    for (int dir : {-s, s}) {
        if (settled) break;
        bool moved = false;
        for (int k = best_k + dir; rc == SID_OK && k >= 0 && k <= ne; k += dir) {
            const double t = eval(k);
            if (rc != SID_OK || t >= 0.99 * best_t) break;
            // confirm against the incumbent timed again after the candidate: one
            // plan's pair time was measured to shift by up to 10 % between
            // builds seconds apart, so a gain must hold on both sides of it
            const double t_again = eval(best_k);
            if (rc != SID_OK || t >= 0.99 * t_again) break;
            best_t = t;
            best_k = k;
            moved = true;
        }
        if (moved || rc != SID_OK) break;
    }
    if (rc == SID_OK) {
        use_prefix(best_k);
        rc = ensure_plan(p, X, Y, nz);
    }
    // The forward's GEMM half is the longer one at the chosen split (its A
    // operand is a volume plane per plane, the backward's the one image), so a
    // larger grid cap for the forward alone may shorten the pair: 1.25 and 1.5
    // blocks per CU are tried on the same terms as the split (at 1005^2 x 51,
    // BF16X2, 26 FFT planes: 1.25 per CU 115.9 RL iter/s against 114.1).
    p->gemm_blocks_f = 0;
    if (rc == SID_OK && !p->gemm_blocks_env && best_k > 0 && best_k < ne) {
        auto timed = [&](int cap) {
            p->gemm_blocks_f = cap;
            const double t = time_projection_pair(p, st, &rc);
            if (log) std::fprintf(stderr, "sid autotune: forward GEMM cap %d blocks: %.3f ms per pair\n",
                                  cap ? cap : p->gemm_blocks, t);
            return t;
        };
        int best_cap = 0;   // 0: the common cap
        double t_inc = timed(best_cap);
        for (int num : {5, 6}) {
            if (rc != SID_OK) break;
            const int cap = (p->gemm_blocks * num / 4 + 7) / 8 * 8;
            const double t = timed(cap);
            if (rc != SID_OK || t >= 0.99 * t_inc) break;
            const double t_again = timed(best_cap);   // the incumbent on both sides, as above
            if (rc != SID_OK || t >= 0.99 * t_again) break;
            best_cap = cap;
            t_inc = t;
        }
        p->gemm_blocks_f = best_cap;
    }
    double d0, d1;
    fft_timing_collect(p, &d0, &d1);  // drop the GEMV events of the trial runs
    if (rc != SID_OK) {
        p->auto_key[0] = -1;  // measure again on the next call
        return rc;
    }
    p->auto_ms[3] = best_k;
    return SID_OK;
}

static int forward_raw(sid_psf *p, const float *V_dev, const int64_t vdims[3], float *O_dev, hipStream_t st,
                       bool sync = true)
{
    SID_TRY(autotune(p, vdims[0], vdims[1], vdims[2], st));
    SID_TRY(ensure_plan(p, vdims[0], vdims[1], vdims[2]));
    Timer tm(p, st);
    const int e0 = tm.mark();
    const unsigned long long stamp = next_stamp(p);
    SID_TRY(launch_vol_in(p, V_dev, vdims[2], 0, st, stamp));
    SID_TRY(project_halves(p, 0, st, tm, 0, 5));
    const int e2 = tm.mark();
    SID_TRY(launch_fwd_reduce(p, O_dev, st, stamp));
    const int e3 = tm.mark();
    tm.span(1, e2, e3);
    if (!sync) return SID_OK;
    SID_TRY_HIP(hipStreamSynchronize(st));
    tm.finish(e0, e3);
    return SID_OK;
}

static int backward_raw(sid_psf *p, const float *O_dev, const int64_t odims[2], float *B_dev, hipStream_t st,
                        bool sync = true)
{
    // the backward volume always has all planes of H (backwardProjectGPU_new.m:4-5); the
    // sensor plan keeps the nz of the last forward call.
    const int64_t nz = (p->plan.vol && p->plan.X == odims[0] && p->plan.Y == odims[1]) ? p->plan.nz : p->nzH;
    SID_TRY(autotune(p, odims[0], odims[1], nz, st));
    SID_TRY(ensure_plan(p, odims[0], odims[1], nz));
    Timer tm(p, st);
    const int e0 = tm.mark();
    const unsigned long long stamp = next_stamp(p);
    SID_TRY(launch_img_in(p, O_dev, st, stamp));
    const BwdTail tail{B_dev, 0, stamp, 0};
    bool direct_done = false, all_done = false;
    SID_TRY(project_halves(p, 1, st, tm, 2, 6, &tail, &direct_done, &all_done));
    const int e2 = tm.mark();
    if (!all_done) SID_TRY(launch_bwd_reduce(p, B_dev, 0, st, stamp, 0, direct_done ? 2 : 0));
    const int e3 = tm.mark();
    tm.span(3, e2, e3);
    if (!sync) return SID_OK;
    SID_TRY_HIP(hipStreamSynchronize(st));
    tm.finish(e0, e3);
    return SID_OK;
}

// Neuron-kernel staging buffer of at least n floats (kept on the handle).
static int kwork_reserve(sid_psf *p, int64_t n)
{
    if (p->kwork_elems >= n) return SID_OK;
    if (p->kwork) (void)hipFree(p->kwork);
    p->kwork = nullptr;
    p->kwork_elems = 0;
    SID_TRY_HIP(hipMalloc(&p->kwork, sizeof(float) * n));
    p->kwork_elems = n;
    return SID_OK;
}

// convn(in, k, 'same') on a handle: the FP32 transform form where it fits (kernel
// spectra kept on the handle; its rounding is below the FP32 direct sum's), else
// the BF16-split MFMA form for BF16X2 / BF16X3 and the FP32 direct sum otherwise
static int handle_convn(sid_psf *p, const float *in, const int64_t vdims[3], float *out, hipStream_t st)
{
    if (convn_fft_enabled() && convn_fft_fits(vdims, p->kdims)) {
        if (!p->kfft) p->kfft = new ConvnFft();
        SID_TRY(convn_fft_prepare(*p->kfft, vdims, p->kern, p->kdims, st));
        return convn_fft_run(*p->kfft, in, out, st);
    }
    if (p->precision == SID_PREC_BF16X2 || p->precision == SID_PREC_BF16X3)
        return convn_mfma(in, vdims, p->kfrag, p->kdims, p->precision, out, st);
    return convn_same(in, vdims, p->kern, p->kdims, out, st);
}

// forwardFUN = F(convn(V, k, 'same')) when the handle carries a neuron kernel (reconstruct_S.m:126,188)
static int forward_impl(sid_psf *p, const float *V_dev, const int64_t vdims[3], float *O_dev, hipStream_t st,
                        bool sync = true)
{
    if (!p->kern) return forward_raw(p, V_dev, vdims, O_dev, st, sync);
    SID_TRY(kwork_reserve(p, vdims[0] * vdims[1] * vdims[2]));
    SID_TRY(handle_convn(p, V_dev, vdims, p->kwork, st));
    return forward_raw(p, p->kwork, vdims, O_dev, st, sync);
}

// backwardFUN = convn(B(O), k, 'same') (reconstruct_S.m:127,189)
static int backward_impl(sid_psf *p, const float *O_dev, const int64_t odims[2], float *B_dev, hipStream_t st,
                         bool sync = true)
{
    if (!p->kern) return backward_raw(p, O_dev, odims, B_dev, st, sync);
    const int64_t vdims[3] = {odims[0], odims[1], p->nzH};
    SID_TRY(kwork_reserve(p, vdims[0] * vdims[1] * vdims[2]));
    SID_TRY(backward_raw(p, O_dev, odims, p->kwork, st, sync));
    SID_TRY(handle_convn(p, p->kwork, vdims, B_dev, st));
    if (sync) SID_TRY_HIP(hipStreamSynchronize(st));
    return SID_OK;
}

int sid_forward_device(sid_psf *p, const float *V, const int64_t vdims[3], float *O, void *stream)
{
    if (!p || !V || !O) { set_error("sid_forward_device: null argument"); return SID_ERR_ARG; }
    SID_TRY(check_vdims(p, vdims, false));
    SID_TRY_HIP(hipSetDevice(p->device));
    hipStream_t st;
    SID_TRY(pick_stream(p, stream, &st));
    return forward_impl(p, V, vdims, O, st);
}

int sid_backward_device(sid_psf *p, const float *O, const int64_t odims[2], float *B, void *stream)
{
    if (!p || !O || !B || !odims) { set_error("sid_backward_device: null argument"); return SID_ERR_ARG; }
    if (odims[0] < 1 || odims[1] < 1) { set_error("image dimensions must be positive"); return SID_ERR_DIMS; }
    SID_TRY_HIP(hipSetDevice(p->device));
    hipStream_t st;
    SID_TRY(pick_stream(p, stream, &st));
    return backward_impl(p, O, odims, B, st);
}

// Host-buffer variants: stage through device buffers kept on the handle
// (hbuf_reserve; a 206 MB hipMalloc + hipFree per call cost more than the
// transfer), or temporary ones (DevBuf) for the optional history.
struct DevBuf {
    void *ptr = nullptr;
    ~DevBuf()
    {
        if (ptr) (void)hipFree(ptr);
    }
};

static int hbuf_reserve(sid_psf *p, int i, int64_t n, float **out)
{
    if (p->hbuf_elems[i] < n) {
        if (p->hbuf[i]) (void)hipFree(p->hbuf[i]);
        p->hbuf[i] = nullptr;
        p->hbuf_elems[i] = 0;
        SID_TRY_HIP(hipMalloc(&p->hbuf[i], sizeof(float) * std::max<int64_t>(n, 1)));
        p->hbuf_elems[i] = n;
    }
    *out = p->hbuf[i];
    return SID_OK;
}

int sid_forward(sid_psf *p, const float *V, const int64_t vdims[3], float *O)
{
    if (!p || !V || !O) { set_error("sid_forward: null argument"); return SID_ERR_ARG; }
    SID_TRY(check_vdims(p, vdims, false));
    SID_TRY_HIP(hipSetDevice(p->device));
    const int64_t nv = vdims[0] * vdims[1] * vdims[2], no = vdims[0] * vdims[1];
    float *dv, *dO;
    SID_TRY(hbuf_reserve(p, 0, nv, &dv));
    SID_TRY(hbuf_reserve(p, 1, no, &dO));
    SID_TRY(copy_h2d(dv, V, nv * sizeof(float), p->stream));
    SID_TRY(forward_impl(p, dv, vdims, dO, p->stream));
    SID_TRY(copy_d2h(O, dO, no * sizeof(float), p->stream));
    return SID_OK;
}

int sid_backward(sid_psf *p, const float *O, const int64_t odims[2], float *B)
{
    if (!p || !O || !B || !odims) { set_error("sid_backward: null argument"); return SID_ERR_ARG; }
    if (odims[0] < 1 || odims[1] < 1) { set_error("image dimensions must be positive"); return SID_ERR_DIMS; }
    SID_TRY_HIP(hipSetDevice(p->device));
    const int64_t no = odims[0] * odims[1], nb = no * p->nzH;
    float *dO, *dB;
    SID_TRY(hbuf_reserve(p, 1, no, &dO));
    SID_TRY(hbuf_reserve(p, 0, nb, &dB));
    SID_TRY(copy_h2d(dO, O, no * sizeof(float), p->stream));
    SID_TRY(backward_impl(p, dO, odims, dB, p->stream));
    SID_TRY(copy_d2h(B, dB, nb * sizeof(float), p->stream));
    SID_TRY_HIP(hipStreamSynchronize(p->stream));
    return SID_OK;
}

static int rl_impl(sid_psf *p, const float *Htf, float *X, const int64_t vdims[3], int max_iter, float *hist,
                   hipStream_t st)
{
    if (p->kern) return rl_wrapped(p, Htf, X, vdims, max_iter, hist, st);
    SID_TRY(autotune(p, vdims[0], vdims[1], vdims[2], st));
    SID_TRY(ensure_plan(p, vdims[0], vdims[1], vdims[2]));
    SensorPlan &sp = p->plan;
    const int64_t nv = vdims[0] * vdims[1] * vdims[2];
    Timer tm(p, st);
    const int e0 = tm.mark();
    // non-finite stamps: the iterate's (set by the staging, then by each update
    // for the next forward) and the image's (set by the forward reduction)
    unsigned long long s_vol = next_stamp(p);
    SID_TRY(launch_vol_in(p, Htf, vdims[2], 1, st));
    SID_TRY(launch_vol_in(p, X, vdims[2], 0, st, s_vol));
    for (int it = 0; it < max_iter; ++it) {
        SID_TRY(project_halves(p, 0, st, tm, 0, 5));
        const int b = tm.mark();
        const unsigned long long s_img = next_stamp(p);
        SID_TRY(launch_fwd_reduce(p, nullptr, st, s_vol, s_img));
        const int c = tm.mark();
        s_vol = next_stamp(p);
        const BwdTail tail{hist ? hist + (int64_t)it * nv : nullptr, 1, s_img, s_vol};
        bool direct_done = false, all_done = false;
        SID_TRY(project_halves(p, 1, st, tm, 2, 6, &tail, &direct_done, &all_done));
        const int d = tm.mark();
        if (!all_done) SID_TRY(launch_bwd_reduce(p, tail.B_col, 1, st, s_img, s_vol, direct_done ? 2 : 0));
        const int e = tm.mark();
        tm.span(1, b, c);
        tm.span(3, d, e);
    }
    SID_TRY(launch_vol_out(p, X, vdims[2], st));
    const int e1 = tm.mark();
    SID_TRY_HIP(hipStreamSynchronize(st));
    tm.finish(e0, e1);
    return SID_OK;
}

int sid_deconv_rl_device(sid_psf *p, const float *Htf, float *X, const int64_t vdims[3], int max_iter, float *hist,
                         void *stream)
{
    if (!p || !Htf || !X) { set_error("sid_deconv_rl_device: null argument"); return SID_ERR_ARG; }
    if (max_iter < 0) { set_error("max_iter must be >= 0"); return SID_ERR_ARG; }
    SID_TRY(check_vdims(p, vdims, true));
    SID_TRY_HIP(hipSetDevice(p->device));
    hipStream_t st;
    SID_TRY(pick_stream(p, stream, &st));
    return rl_impl(p, Htf, X, vdims, max_iter, hist, st);
}

int sid_deconv_rl(sid_psf *p, const float *Htf, float *X, const int64_t vdims[3], int max_iter, float *hist)
{
    if (!p || !Htf || !X) { set_error("sid_deconv_rl: null argument"); return SID_ERR_ARG; }
    if (max_iter < 0) { set_error("max_iter must be >= 0"); return SID_ERR_ARG; }
    SID_TRY(check_vdims(p, vdims, true));
    SID_TRY_HIP(hipSetDevice(p->device));
    const int64_t nv = vdims[0] * vdims[1] * vdims[2];
    DevBuf dhist;
    float *dh, *dx;
    SID_TRY(hbuf_reserve(p, 0, nv, &dh));
    SID_TRY(hbuf_reserve(p, 1, nv, &dx));
    if (hist && max_iter > 0) SID_TRY_HIP(hipMalloc(&dhist.ptr, nv * max_iter * sizeof(float)));
    SID_TRY(copy_h2d(dh, Htf, nv * sizeof(float), p->stream));
    SID_TRY(copy_h2d(dx, X, nv * sizeof(float), p->stream));
    SID_TRY(rl_impl(p, dh, dx, vdims, max_iter, (float *)dhist.ptr, p->stream));
    SID_TRY(copy_d2h(X, dx, nv * sizeof(float), p->stream));
    if (dhist.ptr) SID_TRY(copy_d2h(hist, dhist.ptr, nv * max_iter * sizeof(float), p->stream));
    return SID_OK;
}

static int check_boxes(const sid_psf *p, int64_t ns, const int64_t *boxes, const int64_t sensor[2])
{
    for (int64_t s = 0; s < ns; ++s) {
        const int64_t *b = boxes + 6 * s;
        if (b[0] < 1 || b[1] < b[0] || b[1] > sensor[0] || b[2] < 1 || b[3] < b[2] || b[3] > sensor[1] || b[4] < 1 ||
            b[5] < b[4] || b[5] > p->nzH) {
            set_error("seed %lld: box [%lld %lld %lld %lld %lld %lld] is not inside %lld x %lld x %lld",
                      (long long)s + 1, (long long)b[0], (long long)b[1], (long long)b[2], (long long)b[3],
                      (long long)b[4], (long long)b[5], (long long)sensor[0], (long long)sensor[1],
                      (long long)p->nzH);
            return SID_ERR_DIMS;
        }
    }
    return SID_OK;
}

int sid_library_rects(const sid_psf *p, int64_t ns, const int64_t *boxes, const int64_t sensor[2], int64_t *rects,
                      int64_t *offsets)
{
    if (!p || (ns > 0 && (!boxes || !rects)) || !sensor || !offsets || ns < 0) {
        set_error("sid_library_rects: null argument");
        return SID_ERR_ARG;
    }
    SID_TRY(check_boxes(p, ns, boxes, sensor));
    offsets[0] = 0;
    for (int64_t s = 0; s < ns; ++s) {
        patched_rect(p->n, p->M, sensor[0], sensor[1], boxes + 6 * s, rects + 4 * s);
        const int64_t *r = rects + 4 * s;
        offsets[s + 1] = offsets[s] + (r[1] - r[0] + 1) * (r[3] - r[2] + 1);
    }
    return SID_OK;
}

int sid_library_forward_device(sid_psf *p, int64_t ns, const int64_t *boxes, const float *vals,
                               const int64_t *val_offsets, const int64_t sensor[2], float *out, int flags, void *stream)
{
    if (!p || ns < 0 || (ns > 0 && (!boxes || !vals || !val_offsets || !out)) || !sensor) {
        set_error("sid_library_forward_device: null argument");
        return SID_ERR_ARG;
    }
    if (ns == 0) return SID_OK;
    if (p->kern) {
        set_error("the batched library takes seed volumes already convolved with the neuron kernel "
                  "(convolve with sid_convn_same_device, then project on a handle without a kernel)");
        return SID_ERR_ARG;
    }
    SID_TRY(check_boxes(p, ns, boxes, sensor));
    for (int64_t s = 0; s < ns; ++s) {
        const int64_t *b = boxes + 6 * s;
        const int64_t need = (b[1] - b[0] + 1) * (b[3] - b[2] + 1) * (b[5] - b[4] + 1);
        if (val_offsets[s + 1] - val_offsets[s] != need) {
            set_error("seed %lld: %lld values given, box holds %lld", (long long)s + 1,
                      (long long)(val_offsets[s + 1] - val_offsets[s]), (long long)need);
            return SID_ERR_DIMS;
        }
    }
    SID_TRY_HIP(hipSetDevice(p->device));
    hipStream_t st;
    SID_TRY(pick_stream(p, stream, &st));
    return library_forward(p, ns, boxes, vals, val_offsets, sensor[0], sensor[1], out, st,
                           flags & 1);
}

int sid_library_forward(sid_psf *p, int64_t ns, const int64_t *boxes, const float *vals, const int64_t *val_offsets,
                        const int64_t sensor[2], float *out)
{
    if (!p || ns < 0 || (ns > 0 && (!boxes || !vals || !val_offsets || !out)) || !sensor) {
        set_error("sid_library_forward: null argument");
        return SID_ERR_ARG;
    }
    if (ns == 0) return SID_OK;
    SID_TRY(check_boxes(p, ns, boxes, sensor));
    std::vector<int64_t> rects(4 * ns), offs(ns + 1);
    SID_TRY(sid_library_rects(p, ns, boxes, sensor, rects.data(), offs.data()));
    SID_TRY_HIP(hipSetDevice(p->device));
    DevBuf dv, dout;
    SID_TRY_HIP(hipMalloc(&dv.ptr, sizeof(float) * std::max<int64_t>(val_offsets[ns], 1)));
    SID_TRY_HIP(hipMalloc(&dout.ptr, sizeof(float) * std::max<int64_t>(offs[ns], 1)));
    SID_TRY(copy_h2d(dv.ptr, vals, sizeof(float) * val_offsets[ns], p->stream));
    const int rc = sid_library_forward_device(p, ns, boxes, (const float *)dv.ptr, val_offsets, sensor,
                                              (float *)dout.ptr, 0, p->stream);
    lib_plan_free(p);   // its key holds the buffers freed on return
    SID_TRY(rc);
    return copy_d2h(out, dout.ptr, sizeof(float) * offs[ns], p->stream);
}

int sid_patched_rect(int64_t n, int64_t M, const int64_t sensor[2], const int64_t box[6], int64_t rect[4])
{
    if (!sensor || !box || !rect || n < 1 || M < 1) { set_error("sid_patched_rect: bad argument"); return SID_ERR_ARG; }
    if (box[0] < 1 || box[1] < box[0] || box[2] < 1 || box[3] < box[2]) {
        set_error("sid_patched_rect: empty or non-positive box");
        return SID_ERR_DIMS;
    }
    patched_rect(n, M, sensor[0], sensor[1], box, rect);
    return SID_OK;
}

int sid_library_stats(const sid_psf *p, int64_t st[4])
{
    if (!p || !st) { set_error("null argument"); return SID_ERR_ARG; }
    return library_stats(p, st);
}

static int ls_opts(const sid_ls_opts *o, sid_ls_opts *out)
{
    sid_ls_opts d{};
    d.max_iter = 8;
    if (o) d = *o;
    if (d.max_iter < 0) { set_error("max_iter must be >= 0"); return SID_ERR_ARG; }
    const double l[6] = {d.lamb_TV_L1, d.lamb_TV_L2[0], d.lamb_TV_L2[1], d.lamb_TV_L2[2], d.lamb_L1, d.lamb_L2};
    for (double v : l)
        if (!std::isfinite(v)) { set_error("LS_deconv multipliers must be finite"); return SID_ERR_ARG; }
    *out = d;
    return SID_OK;
}

int sid_ls_deconv_device(sid_psf *p, const float *df, float *X, const int64_t vdims[3], const sid_ls_opts *opts,
                         void *stream)
{
    if (!p || !df || !X) { set_error("sid_ls_deconv_device: null argument"); return SID_ERR_ARG; }
    sid_ls_opts o;
    SID_TRY(ls_opts(opts, &o));
    SID_TRY(check_vdims(p, vdims, true));
    SID_TRY_HIP(hipSetDevice(p->device));
    hipStream_t st;
    SID_TRY(pick_stream(p, stream, &st));
    return ls_deconv(p, df, X, vdims, o, st);
}

int sid_ls_deconv(sid_psf *p, const float *df, float *X, const int64_t vdims[3], const sid_ls_opts *opts)
{
    if (!p || !df || !X) { set_error("sid_ls_deconv: null argument"); return SID_ERR_ARG; }
    sid_ls_opts o;
    SID_TRY(ls_opts(opts, &o));
    SID_TRY(check_vdims(p, vdims, true));
    SID_TRY_HIP(hipSetDevice(p->device));
    const int64_t nv = vdims[0] * vdims[1] * vdims[2];
    float *dh, *dx;
    SID_TRY(hbuf_reserve(p, 0, nv, &dh));
    SID_TRY(hbuf_reserve(p, 1, nv, &dx));
    SID_TRY(copy_h2d(dh, df, nv * sizeof(float), p->stream));
    SID_TRY(ls_deconv(p, dh, dx, vdims, o, p->stream));
    SID_TRY(copy_d2h(X, dx, nv * sizeof(float), p->stream));
    SID_TRY_HIP(hipStreamSynchronize(p->stream));
    return SID_OK;
}

static int check_kdims(const int64_t kd[3])
{
    if (kd[0] < 1 || kd[1] < 1 || kd[2] < 1 || kd[0] > kConvnMaxXY || kd[1] > kConvnMaxXY || kd[2] > kConvnMaxZ) {
        set_error("neuron kernel %lld x %lld x %lld outside 1..%d x 1..%d x 1..%d", (long long)kd[0],
                  (long long)kd[1], (long long)kd[2], kConvnMaxXY, kConvnMaxXY, kConvnMaxZ);
        return SID_ERR_DIMS;
    }
    return SID_OK;
}

int sid_psf_set_kernel(sid_psf *p, const float *k, const int64_t kdims[3])
{
    if (!p) { set_error("null handle"); return SID_ERR_ARG; }
    SID_TRY_HIP(hipSetDevice(p->device));
    if (p->stream) SID_TRY_HIP(hipStreamSynchronize(p->stream));
    if (p->kern) (void)hipFree(p->kern);
    if (p->kfrag) (void)hipFree(p->kfrag);
    if (p->kfft) {
        convn_fft_free(*p->kfft);
        delete p->kfft;
    }
    p->kern = nullptr;
    p->kfrag = nullptr;
    p->kfft = nullptr;
    p->kdims[0] = p->kdims[1] = p->kdims[2] = 0;
    if (!k) return SID_OK;
    if (!kdims) { set_error("sid_psf_set_kernel: null kdims"); return SID_ERR_ARG; }
    SID_TRY(check_kdims(kdims));
    const int64_t nk = kdims[0] * kdims[1] * kdims[2];
    SID_TRY_HIP(hipMalloc(&p->kern, sizeof(float) * nk));
    SID_TRY_HIP(hipMemcpy(p->kern, k, sizeof(float) * nk, hipMemcpyHostToDevice));
    int kd[3];
    for (int d = 0; d < 3; ++d) kd[d] = (int)kdims[d];
    const int rc = convn_build_frag(p->kern, kd, &p->kfrag, p->stream);
    if (rc != SID_OK) {
        (void)hipFree(p->kern);
        p->kern = nullptr;
        return rc;
    }
    SID_TRY_HIP(hipStreamSynchronize(p->stream));
    for (int d = 0; d < 3; ++d) p->kdims[d] = kd[d];
    return SID_OK;
}

// convn on the device: the FP32 transform form where it fits, else the FP32
// direct sum (precision FP32 / FP64) or the BF16-split MFMA form (BF16X2 / BF16X3)
static int convn_dispatch(const float *V, const int64_t vdims[3], const float *k_dev, const int kd[3], int precision,
                          float *out, hipStream_t st)
{
    if (precision != SID_PREC_FP32 && precision != SID_PREC_FP64 && precision != SID_PREC_BF16X2 &&
        precision != SID_PREC_BF16X3) {
        set_error("unknown precision %d", precision);
        return SID_ERR_ARG;
    }
    if (convn_fft_enabled() && convn_fft_fits(vdims, kd)) {
        int dev = 0;
        SID_TRY_HIP(hipGetDevice(&dev));
        return convn_fft_oneshot(dev, V, vdims, k_dev, kd, out, st);
    }
    if (precision == SID_PREC_BF16X2 || precision == SID_PREC_BF16X3) {
        u16 *frag = nullptr;
        int rc = convn_build_frag(k_dev, kd, &frag, st);
        if (rc == SID_OK) rc = convn_mfma(V, vdims, frag, kd, precision, out, st);
        const hipError_t e = hipStreamSynchronize(st);
        if (frag) (void)hipFree(frag);
        if (rc != SID_OK) return rc;
        if (e != hipSuccess) { set_error("convn: %s", hipGetErrorString(e)); return SID_ERR_HIP; }
        return SID_OK;
    }
    SID_TRY(convn_same(V, vdims, k_dev, kd, out, st));
    SID_TRY_HIP(hipStreamSynchronize(st));
    return SID_OK;
}

int sid_convn_same_device(const float *V, const int64_t vdims[3], const float *k, const int64_t kdims[3],
                          int precision, float *out, void *stream)
{
    if (!V || !vdims || !k || !kdims || !out) { set_error("sid_convn_same_device: null argument"); return SID_ERR_ARG; }
    SID_TRY(check_kdims(kdims));
    SID_TRY(check_device());
    const int kd[3] = {(int)kdims[0], (int)kdims[1], (int)kdims[2]};
    return convn_dispatch(V, vdims, k, kd, precision, out, (hipStream_t)stream);
}

int sid_convn_same(const float *V, const int64_t vdims[3], const float *k, const int64_t kdims[3], int precision,
                   float *out)
{
    if (!V || !vdims || !k || !kdims || !out) { set_error("sid_convn_same: null argument"); return SID_ERR_ARG; }
    SID_TRY(check_kdims(kdims));
    if (vdims[0] < 1 || vdims[1] < 1 || vdims[2] < 1) {
        set_error("convn: dimensions must be positive");
        return SID_ERR_DIMS;
    }
    SID_TRY(check_device());
    const int64_t nv = vdims[0] * vdims[1] * vdims[2], nk = kdims[0] * kdims[1] * kdims[2];
    DevBuf dv, dk, dout;
    SID_TRY_HIP(hipMalloc(&dv.ptr, sizeof(float) * nv));
    SID_TRY_HIP(hipMalloc(&dk.ptr, sizeof(float) * nk));
    SID_TRY_HIP(hipMalloc(&dout.ptr, sizeof(float) * nv));
    SID_TRY(copy_h2d(dv.ptr, V, sizeof(float) * nv, nullptr));
    SID_TRY_HIP(hipMemcpy(dk.ptr, k, sizeof(float) * nk, hipMemcpyHostToDevice));
    const int kd[3] = {(int)kdims[0], (int)kdims[1], (int)kdims[2]};
    SID_TRY(convn_dispatch((const float *)dv.ptr, vdims, (const float *)dk.ptr, kd, precision, (float *)dout.ptr,
                           nullptr));
    SID_TRY(copy_d2h(out, dout.ptr, sizeof(float) * nv, nullptr));
    return SID_OK;
}

int sid_library_kernel_boxes(int64_t n_seeds, const int64_t *boxes, const int64_t sensor[2], const int64_t kdims[3],
                             int64_t *boxes_out, int64_t *val_offsets_out)
{
    if (n_seeds < 0 || (n_seeds > 0 && (!boxes || !boxes_out)) || !sensor || !kdims || !val_offsets_out) {
        set_error("sid_library_kernel_boxes: null argument");
        return SID_ERR_ARG;
    }
    SID_TRY(check_kdims(kdims));
    const int kd[3] = {(int)kdims[0], (int)kdims[1], (int)kdims[2]};
    return library_kernel_boxes(n_seeds, boxes, sensor[0], sensor[1], kd, boxes_out, val_offsets_out);
}

int sid_library_convn_seeds_device(int64_t n_seeds, const int64_t *boxes, const float *vals_dev,
                                   const int64_t *val_offsets, const int64_t sensor[2], const float *k_dev,
                                   const int64_t kdims[3], float *vals_out_dev, void *stream)
{
    if (n_seeds < 0 || (n_seeds > 0 && (!boxes || !vals_dev || !val_offsets || !vals_out_dev)) || !sensor || !k_dev ||
        !kdims) {
        set_error("sid_library_convn_seeds_device: null argument");
        return SID_ERR_ARG;
    }
    SID_TRY(check_kdims(kdims));
    SID_TRY(check_device());
    const int kd[3] = {(int)kdims[0], (int)kdims[1], (int)kdims[2]};
    return library_convn_seeds(n_seeds, boxes, vals_dev, val_offsets, sensor[0], sensor[1], k_dev, kd, vals_out_dev,
                               (hipStream_t)stream);
}

int sid_library_convn_seeds(int64_t n_seeds, const int64_t *boxes, const float *vals, const int64_t *val_offsets,
                            const int64_t sensor[2], const float *k, const int64_t kdims[3], float *vals_out)
{
    if (n_seeds < 0 || (n_seeds > 0 && (!boxes || !vals || !val_offsets || !vals_out)) || !sensor || !k || !kdims) {
        set_error("sid_library_convn_seeds: null argument");
        return SID_ERR_ARG;
    }
    SID_TRY(check_kdims(kdims));
    if (n_seeds == 0) return SID_OK;
    std::vector<int64_t> grown(6 * n_seeds), goff(n_seeds + 1);
    const int kd[3] = {(int)kdims[0], (int)kdims[1], (int)kdims[2]};
    SID_TRY(library_kernel_boxes(n_seeds, boxes, sensor[0], sensor[1], kd, grown.data(), goff.data()));
    if (val_offsets[0] != 0) {
        set_error("sid_library_convn_seeds: val_offsets[0] must be 0");
        return SID_ERR_ARG;
    }
    SID_TRY(check_device());
    const int64_t nin = val_offsets[n_seeds], nout = goff[n_seeds], nk = kdims[0] * kdims[1] * kdims[2];
    DevBuf dv, dk, dout;
    SID_TRY_HIP(hipMalloc(&dv.ptr, sizeof(float) * std::max<int64_t>(nin, 1)));
    SID_TRY_HIP(hipMalloc(&dk.ptr, sizeof(float) * nk));
    SID_TRY_HIP(hipMalloc(&dout.ptr, sizeof(float) * std::max<int64_t>(nout, 1)));
    SID_TRY(copy_h2d(dv.ptr, vals, sizeof(float) * nin, nullptr));
    SID_TRY_HIP(hipMemcpy(dk.ptr, k, sizeof(float) * nk, hipMemcpyHostToDevice));
    SID_TRY(library_convn_seeds(n_seeds, boxes, (const float *)dv.ptr, val_offsets, sensor[0], sensor[1],
                                (const float *)dk.ptr, kd, (float *)dout.ptr, nullptr));
    SID_TRY(copy_d2h(vals_out, dout.ptr, sizeof(float) * nout, nullptr));
    return SID_OK;
}

int sid_std_image_device(const float *Y, int64_t P, int64_t T, int bg_iter, int flags, float *s, float *t,
                         float *std_img, double *ymax, void *stream)
{
    if (!Y || !std_img) { set_error("sid_std_image_device: null argument"); return SID_ERR_ARG; }
    if (P < 1 || T < 1 || T > (int64_t)1 << 30) { set_error("movie dimensions out of range"); return SID_ERR_DIMS; }
    if (bg_iter < 0) { set_error("bg_iter must be >= 0"); return SID_ERR_ARG; }
    SID_TRY(check_device());
    DevBuf work;
    SID_TRY_HIP(hipMalloc(&work.ptr, mv_work_bytes(P, T, nullptr, nullptr)));
    return std_image(Y, P, T, bg_iter, flags, s, t, std_img, ymax, (hipStream_t)stream, work.ptr);
}

int sid_std_image_device_noalloc(const float *Y, int64_t P, int64_t T, int bg_iter, int flags, float *s, float *t,
                                 float *std_img, double *ymax, void *stream)
{
    if (!Y || !std_img) { set_error("sid_std_image_device_noalloc: null argument"); return SID_ERR_ARG; }
    if (P < 1 || T < 1 || T > (int64_t)1 << 30) { set_error("movie dimensions out of range"); return SID_ERR_DIMS; }
    if (bg_iter < 0) { set_error("bg_iter must be >= 0"); return SID_ERR_ARG; }
    const size_t need = mv_work_bytes(P, T, nullptr, nullptr);
    if (need > mv_scratch_bytes()) {
        set_error("movie of %lld px x %lld frames needs %zu bytes of work space; the allocation-free entry holds %zu",
                  (long long)P, (long long)T, need, mv_scratch_bytes());
        return SID_ERR_DIMS;
    }
    SID_TRY(check_device());
    return std_image_noalloc(Y, P, T, bg_iter, flags, s, t, std_img, ymax, (hipStream_t)stream);
}

int sid_std_image(const float *Y, int64_t P, int64_t T, int bg_iter, int flags, float *s, float *t, float *std_img,
                  double *ymax)
{
    if (!Y || !std_img) { set_error("sid_std_image: null argument"); return SID_ERR_ARG; }
    if (P < 1 || T < 1 || T > (int64_t)1 << 30) { set_error("movie dimensions out of range"); return SID_ERR_DIMS; }
    if (bg_iter < 0) { set_error("bg_iter must be >= 0"); return SID_ERR_ARG; }
    SID_TRY(check_device());
    DevBuf dy, ds, dt, dstd;
    SID_TRY_HIP(hipMalloc(&dy.ptr, sizeof(float) * P * T));
    SID_TRY_HIP(hipMalloc(&ds.ptr, sizeof(float) * P));
    SID_TRY_HIP(hipMalloc(&dt.ptr, sizeof(float) * T));
    SID_TRY_HIP(hipMalloc(&dstd.ptr, sizeof(float) * P));
    DevBuf work;
    SID_TRY_HIP(hipMalloc(&work.ptr, mv_work_bytes(P, T, nullptr, nullptr)));
    SID_TRY(copy_h2d(dy.ptr, Y, sizeof(float) * P * T, nullptr));
    SID_TRY(std_image((const float *)dy.ptr, P, T, bg_iter, flags, (float *)ds.ptr, (float *)dt.ptr,
                      (float *)dstd.ptr, ymax, nullptr, work.ptr));
    SID_TRY_HIP(hipMemcpy(std_img, dstd.ptr, sizeof(float) * P, hipMemcpyDeviceToHost));
    if (s) SID_TRY_HIP(hipMemcpy(s, ds.ptr, sizeof(float) * P, hipMemcpyDeviceToHost));
    if (t) SID_TRY_HIP(hipMemcpy(t, dt.ptr, sizeof(float) * T, hipMemcpyDeviceToHost));
    return SID_OK;
}

int sid_last_timing(const sid_psf *p, double ms[12])
{
    if (!p || !ms) { set_error("null argument"); return SID_ERR_ARG; }
    for (int i = 0; i < 12; ++i) ms[i] = p->timing[i];
    return SID_OK;
}

int sid_plan_stats(const sid_psf *p, int64_t st[8])
{
    if (!p || !st) { set_error("null argument"); return SID_ERR_ARG; }
    const SensorPlan &sp = p->plan;
    st[0] = sp.fwd.macs; st[1] = sp.bwd.macs;
    st[2] = (int64_t)sp.fwd.items.size(); st[3] = (int64_t)sp.bwd.items.size();
    st[4] = sp.fwd.nslabs; st[5] = sp.bwd.nslabs;
    st[6] = sp.mpad; st[7] = p->NP;
    return SID_OK;
}

int sid_plan_plane_macs(const sid_psf *p, int64_t *fwd, int64_t *bwd, int64_t n)
{
    if (!p || !fwd || !bwd) { set_error("sid_plan_plane_macs: null argument"); return SID_ERR_ARG; }
    if (n != p->nzH) { set_error("sid_plan_plane_macs: n must be size(H,5) = %lld", (long long)p->nzH); return SID_ERR_DIMS; }
    const SensorPlan &sp = p->plan;
    for (int64_t c = 0; c < n; ++c) fwd[c] = bwd[c] = 0;
    if (!sp.vol) { set_error("no sensor plan yet: project once first"); return SID_ERR_STATE; }
    for (size_t i = 0; i < sp.fwd.segs.size(); ++i) fwd[sp.fwd.segs[i].plane] += sp.fwd.cost[i] * p->NP * sp.mpad;
    for (size_t i = 0; i < sp.bwd.segs.size(); ++i) bwd[sp.bwd.segs[i].plane] += sp.bwd.cost[i] * p->NP * sp.mpad;
    return SID_OK;
}

int sid_fft_device(int L, int64_t nseq, float *data, int direction, void *stream)
{
    if (!data || nseq < 0 || (direction != -1 && direction != 1)) { set_error("sid_fft_device: bad argument"); return SID_ERR_ARG; }
    SID_TRY(check_device());
    return fft_test(L, nseq, (float2 *)data, direction, (hipStream_t)stream);
}

int sid_selftest(void)
{
    SID_TRY(check_device());
    return run_selftest();
}

}  // extern "C"

// Projector passes on column-major device buffers for the solvers (checks done by the caller).
int sid::forward_dev(sid_psf *p, const float *V, const int64_t vdims[3], float *O, hipStream_t st, bool sync)
{
    return forward_impl(p, V, vdims, O, st, sync);
}

int sid::backward_dev(sid_psf *p, const float *O, const int64_t odims[2], float *B, hipStream_t st, bool sync)
{
    return backward_impl(p, O, odims, B, st, sync);
}
