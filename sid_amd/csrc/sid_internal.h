// Internal declarations shared by the libsidproj translation units.
//
// Device layout of the projector (lenslet-major, zero-padded):
//   n       pixels per lenslet side (Nnum); Lx = ceil(X/n), Ly = ceil(Y/n)
//   S       per-lenslet stride: n*n entries rounded up to a multiple of 8
//   T       lenslet padding on every side (largest tap radius over planes)
//   XP, YP  padded lenslet grid: XP = Lx + 2T, YP = Ly + 2T + 1
//   volume  vol[c][xp][yp][ab]   ab = a + n*b  (a: row offset, b: column offset in the lenslet)
//   image   img[xp][yp][uv]      uv = u + n*v
// The projector is a sum of GEMMs over lenslet tap rows dx (see sid_direct.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/sid_proj.h"

namespace sid {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef double d4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

// FP32 -> BF16 bits (round to nearest even, NaN kept) and back.
// FP32 -> BF16, round to nearest even (one v_cvt_pk_bf16_f32 on gfx950; NaN
// stays NaN, infinities and overflow to infinity as IEEE rounding gives)
__device__ __forceinline__ u16 f2bf(float x) { return __builtin_bit_cast(u16, (__bf16)x); }
__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float(((uint32_t)h) << 16); }

void set_error(const char *fmt, ...);

#define SID_TRY_HIP(call)                                                                    \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) {                                                              \
            ::sid::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                             __LINE__);                                                      \
            return e_ == hipErrorOutOfMemory ? SID_ERR_OOM : SID_ERR_HIP;                    \
        }                                                                                    \
    } while (0)

#define SID_TRY(expr)               \
    do {                            \
        int s_ = (expr);            \
        if (s_ != SID_OK) return s_; \
    } while (0)

// One GEMM segment: a (plane, tap row) pair.  For output row r (lenslet
// (Px, Py) = (r / Ly, r % Ly)) the A operand row starts at
// a_off + (Px*YP + Py)*S and runs K contiguous elements (Toeplitz along y);
// the B operand is K x NP contiguous at w_off.
struct Seg {
    int64_t a_off;
    int64_t w_off;     // FP32 / FP64 weights: K = taps * S padded to KPAD
    int32_t K;
    int32_t plane;
    int64_t w_off2;    // BF16 weights: K2 = taps * KT (each tap padded to KT = S rounded up to 32)
    int32_t K2;
    int32_t tmap;      // BF16: offset (32-bit words) of the segment's zero-tile table
};

// One block of work: segments [seg0, seg1) for M tile `mtile`, written to
// partial-sum slab `slab`.
struct Item {
    int32_t seg0, seg1, mtile, slab;
};

// Per-plane kernel support and lenslet tap ranges.
struct Plane {
    int lo, hi;          // inclusive kernel index range used (both axes)
    int off_f, off_b;    // p - q = i - off (forward / backward)
    int tf_lo, tf_hi;    // forward taps dx, dy in [tf_lo, tf_hi]
    int tb_lo, tb_hi;    // backward taps
    int64_t Kf, Kb;      // padded GEMM K per tap row
    int64_t wf_off, wb_off;  // weight offsets (elements)
    int64_t Kf2, Kb2;        // BF16 layout: taps * KT
    int64_t wf2_off, wb2_off;
    bool empty;
};

constexpr int KPAD = 64;     // K is padded to a multiple of this
constexpr int NTHREADS = 256;

struct GemmPlan {             // one direction on one sensor geometry
    std::vector<Seg> segs;
    std::vector<Item> items;
    std::vector<int> slab_plane;      // backward: plane of each slab
    std::vector<int> plane_slab0;     // backward: first slab of plane c (size nz+1)
    Seg *d_segs = nullptr;
    Item *d_items = nullptr;
    int *d_plane_slab0 = nullptr;
    std::vector<uint32_t> tmap;       // zero-tile tables of the BF16 segments (Seg::tmap)
    std::vector<int64_t> cost;        // per segment: K rows the GEMM runs at the plan's precision
    uint32_t *d_tmap = nullptr;
    std::vector<int64_t> steps;       // BF16 GEMM step table: per step (A offset, B offset << 1 | continues chunk)
    std::vector<int32_t> run0;        // first step of slab i's run of segments (size nslabs + 1)
    int64_t *d_steps = nullptr;
    int32_t *d_run0 = nullptr;
    int nslabs = 0;
    int64_t macs = 0;                 // MFMA multiply-adds issued (padding included)
};

struct FftPlan1D {
    int L = 0, nst = 0;
    int rad[12] = {};
    float2 *tw_fwd = nullptr, *tw_inv = nullptr;
};

// FFT path state for one sensor geometry (sid_fft.hip).
// One group of the tiled FFT path (sid_tfft.hip).
struct TfGroup {
    int s0 = 0, ns = 0;                // slots [s0, s0 + ns) of FftState::planes
    int T = 0, H = 0, B = 0;           // FFT length, stored half length T / 2 + 1, valid outputs per tile side
    int Dlo = 0, Dhi = 0, K = 0;       // common lenslet tap range of the group's planes and its width
    int ntx = 0, nty = 0, ntile = 0, nbin = 0, kpad = 0;
    int tsplit_f = 1, tchunk_f = 0;    // forward GEMM: K-step runs per bin and their length
    int tsplit_b = 1, tchunk_b = 0;    // backward GEMM: slot runs per bin and their length
    FftPlan1D ft;                      // twiddles of length T (spectra build)
    void *G = nullptr, *Gb = nullptr;  // forward / backward spectra, MFMA fragment order
    float2 *vin = nullptr;             // forward input spectra [nbin][kpad / 16][ntile][16] (K padding stays zero)
    float2 *zin = nullptr;             // backward input (image) spectra [nbin][16][ntile][16] (uv padding zero)
    float *vmax = nullptr;             // per tile and transform block: bound on its spectra
    float *tmax = nullptr;             // per tile: bound over all its blocks
    int chain = 0;                     // 0: the caller's stream, 1: the handle's second stream
    double cost = 0;                   // modelled time (s) of one forward + backward
};

struct FftState {
    bool active = false;
    std::vector<int> planes;       // FFT planes, in slot order
    std::vector<int> slot_of;      // per plane of H: slot or -1
    int Nxp = 0, Nyp = 0, Nx = 0, Ny = 0, Hy = 0, nk0 = 0, n2 = 0, nq = 0;
    int nsplit = 1, qc = 0;        // forward GEMV row chunks per k0 and rows per chunk
    int nsplit_b = 1, qc_b = 0;    // the same for the backward GEMV
    bool fused = false;            // sub-image transforms in one LDS-resident kernel per pair
    size_t sub_lds = 0;
    int sub_blocks = 0;            // persistent grid of the fused kernels (one block per CU)
    FftPlan1D fx, fy, fX, fY;      // lengths Nx', Ny', Nx, Ny
    void *G = nullptr;             // spectra [nk0][nq][n2] (k0 blocks kstride elements apart):
                                   // float2, or __half2 scaled by 1/qscale[q] when `half`
    bool half = false;
    int64_t kstride = 0;
    float *qscale = nullptr;       // per row q: max |G| bound (L1 norm of the kernel), half only
    float2 *subs = nullptr;        // sub-image transforms [nq][Nx'][Ny'] (unfused path only)
    float2 *vt = nullptr;          // [nk0][nq]
    float2 *oh = nullptr;          // forward GEMV chunk sums [nsplit][nk0][n2]
    float2 *ohT = nullptr;         // their sum, transposed [n2][nk0]
    float2 *grid = nullptr;        // [Nx][Ny]
    float *dfft = nullptr;         // backward result of the FFT planes, padded volume layout
    int *d_slot_plane = nullptr;
    int *d_plane_fft = nullptr;    // per plane of H: 1 when on the FFT path
    int *d_plane_list = nullptr;   // planes of H on the direct GEMM, then those on the FFT path
    int n_list_direct = 0, n_list_fft = 0;
    bool all_fft = false;           // every nonempty plane of H is on the FFT path
    int64_t g_bytes = 0;
    // Tiled form (BF16X2, sid_tfft.hip): the FFT planes in groups of similar
    // lenslet tap range (slots in ascending tap width); per group overlap-save
    // tiles of T x T lenslets over the lenslet sub-images and, per frequency
    // bin, one complex GEMM over all tiles (FP16 spectra x FP16 hi + lo inputs
    // on MFMA, FP32 accumulate).
    bool tiled = false;
    std::vector<TfGroup> tg;
    int chp = 0, kslot = 0, ngrp = 0;  // GEMM rows (n^2 padded to 16), K per slot (n^2 padded to 4), channel groups
    float2 *yout = nullptr;            // forward GEMM partial results, largest group [tsplit_f][nbin][ntp][chp]
    float2 *bout = nullptr;            // backward GEMM result, largest group [nbin][slot][ntp][chp]
    float *tout = nullptr;             // forward result of the FFT planes, lenslet image [nrows][NP]
    // With two chains (every plane on the FFT path, two or more groups), the
    // groups of chain 1 run on the handle's second stream beside those of chain
    // 0, each chain with its own GEMM results and forward image.
    int nchain = 1;
    float2 *yout2 = nullptr, *bout2 = nullptr;
    float *tout2 = nullptr;
    hipEvent_t tev_fork = nullptr, tev_join = nullptr;
    std::vector<hipEvent_t> ev;     // event pairs bracketing the GEMVs of the last call
    std::vector<int> ev_kind;        // 0 forward GEMV, 1 backward GEMV (one entry per pair)
};

struct SensorPlan {
    int64_t X = 0, Y = 0, nz = 0;     // nz of the forward input
    int64_t Lx = 0, Ly = 0, XP = 0, YP = 0;
    int64_t nrows = 0, ntiles = 0, mpad = 0, mt = 0;
    GemmPlan fwd, bwd;
    int cost_bf16 = -1;               // segment costs of the BF16 GEMM (tap hulls) or of K
    // device buffers (fp32 master + bf16 hi/lo shadows)
    float *vol = nullptr, *img = nullptr, *htf = nullptr, *part = nullptr;
    u16 *vol_hi = nullptr, *vol_lo = nullptr, *img_hi = nullptr, *img_lo = nullptr;
    double *vol_d = nullptr, *img_d = nullptr, *htf_d = nullptr, *part_d = nullptr;  // FP64 mode
    int64_t vol_elems = 0, img_elems = 0, part_elems = 0;
    int prec_built = -1;
    FftState fft;
    std::vector<char> fft_use;       // per plane of H: on the FFT path in this plan
};

struct ConvnFft;   // sid_convn_fft.hip

}  // namespace sid

struct sid_psf {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t M = 0, n = 0, nzH = 0;
    int64_t S = 0, NP = 0, T = 0;
    bool cropped = false;
    bool full_support = false;
    std::vector<sid::Plane> planes;
    float *H = nullptr;               // device copy, MATLAB layout
    float *wf = nullptr, *wb = nullptr;           // fp32 weights (forward / backward)
    sid::u16 *wf_hi = nullptr, *wb_hi = nullptr;  // bf16 hi weights
    sid::u16 *wf_lo = nullptr, *wb_lo = nullptr;  // bf16 lo weights (x3 mode)
    double *wf_d = nullptr, *wb_d = nullptr;      // fp64 weights
    int64_t wf_elems = 0, wb_elems = 0;
    int64_t wf2_elems = 0, wb2_elems = 0;   // BF16 weights (tap-padded layout)
    int64_t KT = 0;                         // BF16 K per tap (S rounded up to 32)
    // per 32 x 16 tile of the BF16 weight layout (512 consecutive entries):
    // 1 when an entry is nonzero (forward / backward)
    std::vector<uint8_t> tflag_f, tflag_b;
    // per 32-row chunk of the FP32 weight layout (NP * 32 consecutive entries)
    std::vector<uint8_t> cflag_f, cflag_b;
    int precision = SID_PREC_FP32;
    int method = SID_METHOD_AUTO;
    std::vector<char> pin;            // per plane: 0 autotuned, SID_METHOD_DIRECT / _FFT pinned by the caller
    std::vector<char> auto_fft;       // autotuned per-plane FFT choice (SID_METHOD_AUTO)
    int64_t auto_key[4] = {-1, -1, -1, -1};  // X, Y, nz, precision the choice was measured for
    double auto_ms[4] = {0, 0, 0, 0};        // direct ms, fft ms (all eligible planes), n eligible, n chosen
    bool auto_refined = false;               // the choice has been re-timed with the halves overlapped
    sid::SensorPlan plan;
    void *lib = nullptr;              // cached library plan (sid_library.hip)
    float *kern = nullptr;            // neuron kernel (device, [kx][ky][kz]); null = none
    sid::u16 *kfrag = nullptr;        // its Toeplitz B fragments for the BF16 MFMA convn
    sid::ConvnFft *kfft = nullptr;    // its transform form, built at the first convn
    int kdims[3] = {0, 0, 0};
    float *kwork = nullptr;           // convn staging volume (column-major), kwork_elems floats
    float *hbuf[2] = {nullptr, nullptr};   // device copies of the host-buffer entries' arrays
    int64_t hbuf_elems[2] = {0, 0};
    int64_t kwork_elems = 0;
    hipEvent_t ev[10] = {};
    double timing[12] = {0};
    // second stream for the direct GEMM while the FFT path runs on the caller's
    // stream (fork / join events); off when SID_NO_OVERLAP is set at creation
    bool overlap = true;
    int gemm_blocks = 0;              // grid cap of the GEMM beside the FFT path (0: one block per CU)
    int gemm_blocks_f = 0;            // the forward's cap when the autotuner chose another one (0: gemm_blocks)
    bool gemm_blocks_env = false;     // the cap was set with SID_GEMM_BLOCKS (not tuned)
    hipStream_t aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipEvent_t ev_null = nullptr;     // orders the handle's stream after the default stream (stream = NULL)
    // Non-finite inputs (forwardProjectGPU's rule, conv2FFT_.m:14: one NaN or Inf
    // anywhere in a projection's input makes the whole output NaN, for every
    // plane method).  nf[0] / nf[1]: the stamp of the last forward / backward
    // input that held a non-finite value.  Every staged input gets a fresh stamp
    // from nf_epoch (never reused), so the flags are never cleared.
    unsigned long long *nf = nullptr;
    unsigned long long nf_epoch = 0;
};

namespace sid {
// sid_direct.hip
int build_weights(sid_psf *p, int which /*0 fp32, 1 bf16 hi, 2 bf16 lo, 3 fp64*/);
int ensure_plan(sid_psf *p, int64_t X, int64_t Y, int64_t nz);
int ensure_precision_buffers(sid_psf *p);
// max_blocks > 0 caps the grid (blocks then loop over the items)
int launch_gemm(sid_psf *p, int dir, hipStream_t st, int max_blocks = 0);
// Non-finite stamps (sid_psf::nf): `mark` != 0 stamps the flag the kernel's
// output feeds when it sees a non-finite value (staging: its input); `check`
// != 0 makes the reduction write NaN to every sensor pixel / voxel when the
// input's flag carries that stamp.  0 turns either off.
int launch_fwd_reduce(sid_psf *p, float *O_col, hipStream_t st, unsigned long long check = 0,
                      unsigned long long mark = 0);
int launch_bwd_reduce(sid_psf *p, float *B_col, int rl_update, hipStream_t st, unsigned long long check = 0,
                      unsigned long long mark = 0, int subset = 0);
int launch_vol_in(sid_psf *p, const float *V_col, int64_t nz, int which /*0 iterate, 1 Htf*/, hipStream_t st,
                  unsigned long long mark = 0);
int launch_img_in(sid_psf *p, const float *O_col, hipStream_t st, unsigned long long mark = 0);
inline unsigned long long next_stamp(sid_psf *p) { return ++p->nf_epoch; }
int launch_vol_out(sid_psf *p, float *V_col, int64_t nz, hipStream_t st);
void free_plan_buffers(SensorPlan &sp);
int support_scan(sid_psf *p);
void free_weights(sid_psf *p);
// sid_fft.hip
int fft_plan_make(FftPlan1D &pl, int L);
void fft_plan_free(FftPlan1D &pl);
int fft_run(const FftPlan1D &pl, float2 *data, int64_t nseq, int64_t inner, int64_t outer_stride, int64_t seq_stride,
            int64_t elem_stride, int dir, float scale, hipStream_t st);
void fft_state_free(FftState &f);
bool fft_eligible(const sid_psf *p, int64_t c);
int fft_state_build(sid_psf *p, const std::vector<char> &use, hipStream_t st);
int fft_forward(sid_psf *p, hipStream_t st);
// The backward reduction that follows a backward projection (the arguments
// of launch_bwd_reduce): the column-major output (or null), the ISRA update
// flag, and the non-finite stamps to check and to set.
struct BwdTail {
    float *B_col;
    int rl_update;
    unsigned long long check, mark;
};
int fft_backward(sid_psf *p, hipStream_t st, const BwdTail *tail = nullptr, bool *tail_done = nullptr);
void fft_timing_collect(sid_psf *p, double *fwd_ms, double *bwd_ms);
int next_smooth235(int L);
int fft_test(int L, int64_t nseq, float2 *data, int dir, hipStream_t st);
int autotune_methods(sid_psf *p, int64_t X, int64_t Y, int64_t nz, hipStream_t st);
// sid_tfft.hip: the tiled form of the FFT path.  tf_build returns
// kTfNotApplicable (nothing allocated) when the geometry has no tile size.
constexpr int kTfNotApplicable = 1000;
int tf_build(sid_psf *p, hipStream_t st);
void tf_free(FftState &f);
// sid_api.hip: the handle's second stream and its fork / join events (created on first use)
int ensure_aux(sid_psf *p);
// bytes (at least) and issued MFMA flop of one forward / backward projection's spectra GEMMs
void tf_gemm_work(const FftState &f, int64_t bytes[2], double flops[2]);
int tf_forward(sid_psf *p, hipStream_t st);    // FFT planes' forward image into f.tout
// FFT planes' back-projection into f.dfft, or through the reduction's tail (see BwdTail)
int tf_backward(sid_psf *p, hipStream_t st, const BwdTail *tail = nullptr, bool *tail_done = nullptr);
void fft_mark(FftState &f, hipStream_t st, int kind, bool begin);   // timing events (GEMV / GEMM)
// sid_library.hip
void patched_rect(int64_t n, int64_t M, int64_t X, int64_t Y, const int64_t box[6], int64_t rect[4]);
int library_forward(sid_psf *p, int64_t ns, const int64_t *boxes, const float *vals, const int64_t *val_off, int64_t X,
                    int64_t Y, float *out, hipStream_t st, int use_graph);
void lib_plan_free(sid_psf *p);
int library_stats(const sid_psf *p, int64_t st[4]);
int library_kernel_boxes(int64_t ns, const int64_t *boxes, int64_t X, int64_t Y, const int kd[3], int64_t *out,
                         int64_t *voff_out);
int library_convn_seeds(int64_t ns, const int64_t *boxes, const float *vals, const int64_t *voff, int64_t X,
                        int64_t Y, const float *k_dev, const int kd[3], float *out, hipStream_t st);
// sid_grad.hip
int grad_core(sid_psf *p, const float *V, const float *y, const int64_t dims[4], int mode, double *nll, float *gV,
              float *gH, hipStream_t st);
int run_selftest();
// FP64 B(1) = backward(ones(X, Y)) on the handle's support crops, column-major [X, Y, Nz] (device)
int backward_ones_f64(sid_psf *p, int64_t X, int64_t Y, double *B1, hipStream_t st);
}  // namespace sid

namespace sid {
// sid_api.hip: projector passes on column-major device buffers (used by the solvers)
// sid_hostio.hip: large pageable-host <-> device copies through pinned slots on
// several host threads.  copy_h2d returns once the source has been read (the
// DMAs are ordered before later work on st); copy_d2h returns with dst filled.
int copy_h2d(void *dst_dev, const void *src_host, size_t bytes, hipStream_t st);
int copy_d2h(void *dst_host, const void *src_dev, size_t bytes, hipStream_t st);
// sync = false leaves the work queued on st (no host wait, no per-call timing;
// the caller drains the FFT path's timing events with fft_timing_collect)
int forward_dev(sid_psf *p, const float *V, const int64_t vdims[3], float *O, hipStream_t st, bool sync = true);
int backward_dev(sid_psf *p, const float *O, const int64_t odims[2], float *B, hipStream_t st, bool sync = true);
// sid_convn.hip: convn(in, k, 'same') on column-major device volumes (k on the device, [kx][ky][kz])
constexpr int kConvnMaxXY = 65, kConvnMaxZ = 255;
int convn_same(const float *in, const int64_t vdims[3], const float *k_dev, const int kd[3], float *out,
               hipStream_t st);
// BF16-split MFMA form (precision BF16X2 / BF16X3) with Toeplitz B fragments built once per kernel
int convn_build_frag(const float *k_dev, const int kd[3], u16 **frag, hipStream_t st);
int convn_mfma(const float *in, const int64_t vdims[3], const u16 *frag, const int kd[3], int precision, float *out,
               hipStream_t st);
// sid_convn_fft.hip: the same convolution by 2-D transforms in (x, y) and a
// direct sum along z between them (FP32 arithmetic, every precision where it fits)
struct ConvnFft {
    int X = 0, Y = 0, Z = 0, kx = 0, ky = 0, kz = 0;
    int Nx = 0, Ny = 0, Hx = 0;           // transform lengths, stored x bins Nx / 2 + 1
    int Yp = 0, Pc = 0;                   // rows the x passes touch (Y up to 16), column pitch
    FftPlan1D px, py;
    float2 *kf = nullptr;                 // kernel spectra [kz][Hx][Ny], scaled by 1 / (Nx Ny)
    float2 *s1 = nullptr, *s2 = nullptr;  // spectra by column, in / out [Z][Hx][Pc]
    size_t s_elems = 0;                   // capacity of s1 and s2 (complex values each)
    bool kf_ready = false;
};
// the transform form applies to this volume and kernel (sizes, kz <= 8)
bool convn_fft_fits(const int64_t vdims[3], const int kd[3]);
// (re)size c for the geometry and transform the kernel (when kf_ready is false)
int convn_fft_prepare(ConvnFft &c, const int64_t vdims[3], const float *k_dev, const int kd[3], hipStream_t st);
int convn_fft_run(ConvnFft &c, const float *in, float *out, hipStream_t st);
void convn_fft_free(ConvnFft &c);
// SID_CONVN_FFT=0 keeps convn on the direct forms (FP32 sum, BF16 MFMA)
bool convn_fft_enabled();
// handle-free convn on `device` with its per-device work arrays (kept for the
// process); synchronises st
int convn_fft_oneshot(int device, const float *in, const int64_t vdims[3], const float *k_dev, const int kd[3],
                      float *out, hipStream_t st);
// sid_movie.hip: movie background (rank 1) and std image
// work: mv_work_bytes(P, T, nullptr, nullptr) bytes of device memory
struct MvWork;
size_t mv_work_bytes(int64_t P, int64_t T, MvWork *w, void *base);
int std_image(const float *Y, int64_t P, int64_t T, int bg_iter, int flags, float *s_out, float *t_out,
              float *std_img, double *ymax_out, hipStream_t st, void *work);
// the same with the work space in the library's per-device scratch array (no allocation)
int std_image_noalloc(const float *Y, int64_t P, int64_t T, int bg_iter, int flags, float *s_out, float *t_out,
                      float *std_img, double *ymax_out, hipStream_t st);
size_t mv_scratch_bytes();
// sid_solvers.hip
int ls_deconv(sid_psf *p, const float *df, float *X, const int64_t vdims[3], const sid_ls_opts &o, hipStream_t st);
// deconvRL basic on column-major device buffers through forward_dev / backward_dev (neuron-kernel handles)
int rl_wrapped(sid_psf *p, const float *Htf, float *X, const int64_t vdims[3], int max_iter, float *hist,
               hipStream_t st);
}  // namespace sid
