/*
 * sid_proj.h — C ABI of the MI355X (gfx950) light-field projection library
 * (libsidproj.so).  Replaces SID's MATLAB projectors and ISRA loop:
 *
 *   forwardProjectGPU(H, realspace)          external/LFrecon_util/forwardProjectGPU.m:1
 *   backwardProjectGPU_new(H, projection)    external/LFrecon_util/backwardProjectGPU_new.m:1
 *   forwardProjectACC(H, realspace, CAindex) external/LFrecon_util/forwardProjectACC.m:1
 *   backwardProjectACC_new(H, proj, CAindex) external/LFrecon_util/backwardProjectACC_new.m:1
 *   deconvRL(F, B, Htf, maxIter, Xguess)     external/LFrecon_util/deconvRL.m:1 (basic, lambda 0)
 *
 * Layout contract: every array is MATLAB column-major, exactly as a MEX
 * gateway receives it (mxGetData):
 *   H        [M, M, Nnum, Nnum, Nz]  single   H(i,j,aa,bb,cc) at i + M*(j + M*(aa + Nnum*(bb + Nnum*cc)))
 *   volume   [X, Y, nz]              single   V(x,y,c) at x + X*(y + Y*c)
 *   image    [X, Y]                  single   O(x,y) at x + X*y
 *   CAindex  [Nz, 2]                 int32    1-based inclusive crop, CAindex(c,k) at c + Nz*(k-1)
 *
 * Ownership: the caller owns every input and output buffer.  The library owns
 * the device-resident PSF state behind an opaque sid_psf handle (H is uploaded
 * once per handle, not once per call as forwardProjectGPU.m:19 does).
 *
 * Errors: every entry point returns a sid_status; sid_last_error() gives the
 * message of the last failure on the calling thread.  A handle is used by one
 * thread at a time.  All calls are synchronous: when one returns, its outputs
 * are written (device outputs included).
 *
 * "stream" arguments take a hipStream_t cast to void*; device-pointer entry
 * points read and write device memory in stream order.  NULL is the device's
 * default (null) stream: the call's work is ordered after everything already
 * queued there (the handle runs it on its own stream after an event wait).
 */
#ifndef SID_PROJ_H
#define SID_PROJ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sid_psf sid_psf;

typedef enum {
    SID_OK = 0,
    SID_ERR_ARG = 1,        /* null pointer, bad option */
    SID_ERR_DIMS = 2,       /* inconsistent or unsupported dimensions */
    SID_ERR_NZ = 3,         /* volume has more planes than H (forwardProjectGPU.m:18-19) */
    SID_ERR_OOM = 4,        /* device allocation failed */
    SID_ERR_HIP = 5,        /* HIP runtime error */
    SID_ERR_STATE = 6,      /* call not valid in the handle's state */
    SID_ERR_NODEV = 7       /* no usable gfx950 device */
} sid_status;

/* Arithmetic of the projector GEMMs. */
typedef enum {
    SID_PREC_FP32 = 0,      /* plain method: FP32 operands, FP32 MFMA accumulate */
    SID_PREC_BF16X2 = 1,    /* PSF in BF16, volume/image split hi+lo BF16, FP32 accumulate */
    SID_PREC_BF16X3 = 2,    /* PSF and volume/image both split hi+lo (3 products), FP32 accumulate */
    SID_PREC_FP64 = 3       /* FP64 operands and MFMA (reference path for error studies); float I/O */
} sid_precision;

/* Per-plane method of the projector. */
typedef enum {
    SID_METHOD_AUTO = 0,    /* per-plane choice from the autotuner */
    SID_METHOD_DIRECT = 1,  /* implicit GEMM over lenslet taps on MFMA */
    SID_METHOD_FFT = 2      /* sublattice-periodised spectra, radix-2/3/5 FFT */
} sid_method;

const char *sid_last_error(void);
int sid_version(void);
/* Number of visible devices that are gfx950; 0 when none (no error). */
int sid_device_count(void);

/* --- PSF handle ---------------------------------------------------------- */

/* Upload H (host memory) and build the device-side projector state.
 * CAindex: NULL reproduces forwardProjectGPU / backwardProjectGPU_new (the full
 * M x M kernel, forwardProjectGPU.m:19); non-NULL reproduces forwardProjectACC /
 * backwardProjectACC_new (kernel cropped to CAindex(cc,1):CAindex(cc,2),
 * forwardProjectACC.m:22, backwardProjectACC_new.m:21). */
int sid_psf_create(const float *H, const int64_t dims[5], const int32_t *CAindex, sid_psf **out);
/* Same with H already in device memory (same layout); H is copied. */
int sid_psf_create_device(const float *H_dev, const int64_t dims[5], const int32_t *CAindex,
                          sid_psf **out);
int sid_psf_destroy(sid_psf *psf);
/* SID_OK when psf is a handle created in this process and not yet destroyed,
 * else SID_ERR_ARG.  No counterpart in the reference; for callers that carry
 * handles through integers (a MEX gateway's uint64 handle argument). */
int sid_psf_is_live(const sid_psf *psf);

/* Creation flags for sid_psf_create_ex. */
enum {
    SID_FLAG_H_ON_DEVICE = 1,   /* H points to device memory */
    SID_FLAG_FULL_SUPPORT = 2   /* keep every kernel entry of the M x M (or CAindex) box instead of
                                   shrinking each plane to its nonzero support; needed when H is
                                   later updated with a wider support (PSF refinement) */
};
int sid_psf_create_ex(const float *H, const int64_t dims[5], const int32_t *CAindex, int flags,
                      sid_psf **out);
/* Replace the values of H (same dims; host or device memory per `on_device`)
 * and rebuild every weight copy the handle holds (the sensor plan, with its
 * FFT spectra, is built again by the next projection).  The per-plane support
 * chosen at creation is kept: entries outside it are ignored. */
int sid_psf_update(sid_psf *psf, const float *H, int on_device);

/* Select the GEMM precision (builds the BF16 weight copies on first use). */
int sid_psf_set_precision(sid_psf *psf, int precision);
/* Force a method for every plane (SID_METHOD_DIRECT / _FFT) or restore the
 * autotuned per-plane choice (SID_METHOD_AUTO). */
int sid_psf_set_method(sid_psf *psf, int method);
/* Use the handle's second stream (on != 0, the default unless SID_NO_OVERLAP=1
 * was set at creation) or only the caller's stream: the direct GEMM beside the
 * FFT path, and the two chains of tiled FFT plane groups.  Drops the sensor
 * plan (the next projection builds it again).  No counterpart in the
 * reference; a measurement control (kernel times without concurrency). */
int sid_psf_set_overlap(sid_psf *psf, int on);
/* Pin the method of single planes under SID_METHOD_AUTO: methods[c] is
 * SID_METHOD_DIRECT or SID_METHOD_FFT for plane c, or SID_METHOD_AUTO to leave
 * that plane to the autotuner; n = size(H,5).  methods = NULL clears every pin.
 * A plane the FFT path cannot take (empty, or an even kernel support) is
 * rejected with SID_ERR_ARG.  No counterpart in the reference, whose
 * forwardProjectGPU.m convolves every plane by FFT; for callers that fix the
 * plan of a workload they have already tuned. */
int sid_psf_set_plane_methods(sid_psf *psf, const int32_t *methods, int64_t n);

typedef struct {
    int64_t M, Nnum, Nz;
    int64_t taps_max;          /* largest lenslet tap radius over planes */
    int64_t weight_bytes;      /* device bytes held by projector weights */
    int64_t support_lo[512];   /* per-plane inclusive kernel support used (0-based) */
    int64_t support_hi[512];
    int32_t plane_method[512]; /* per plane in the current sensor plan: SID_METHOD_DIRECT / _FFT, 0 = empty */
    double autotune_ms[4];     /* last autotune: direct ms, FFT ms (all eligible planes), #eligible, #chosen */
    int64_t spectra_bytes;     /* device bytes of FFT-path spectra in the current plan */
    int32_t fft_tile[4];       /* FFT path in tiles, its widest plane group: FFT length T, valid lenslets
                                  per tile side, tiles, common lenslet taps K; all 0 for spectra of the
                                  whole padded sensor */
    int32_t fft_groups;        /* plane groups of the tiled FFT path (0 when not tiled) */
    int32_t fft_group_planes[8]; /* planes per group, narrowest taps first (first 8 groups) */
    int32_t fft_group_T[8];    /* FFT length per group */
    int64_t fft_gemm_bytes[2]; /* tiled path, all groups: bytes one forward / backward projection's spectra
                                  GEMMs must move at least (FP16 spectra + input spectra read + results
                                  written) */
    double fft_gemm_flops[2];  /* the same GEMMs' MFMA work as issued (padded rows, tiles and K; FP16 hi
                                  and lo products), flop */
} sid_psf_info;
int sid_psf_get_info(const sid_psf *psf, sid_psf_info *info);

/* --- projectors (host buffers: the MEX path) ------------------------------ */

/* forwardProjectGPU.m:1 / forwardProjectACC.m:1.  vdims = {X, Y, nz},
 * nz <= Nz of H; O is X*Y. */
int sid_forward(sid_psf *psf, const float *V, const int64_t vdims[3], float *O);
/* backwardProjectGPU_new.m:1 / backwardProjectACC_new.m:1.  odims = {X, Y};
 * B is X*Y*Nz (all planes of H, backwardProjectGPU_new.m:4-5). */
int sid_backward(sid_psf *psf, const float *O, const int64_t odims[2], float *B);

/* --- projectors (device buffers) ----------------------------------------- */
int sid_forward_device(sid_psf *psf, const float *V_dev, const int64_t vdims[3], float *O_dev,
                       void *stream);
int sid_backward_device(sid_psf *psf, const float *O_dev, const int64_t odims[2], float *B_dev,
                        void *stream);

/* --- ISRA / Richardson-Lucy loop (deconvRL.m:46-93, basic, lambda = 0) ----
 * Runs all iterations on the device.  X holds Xguess on entry and the
 * result on exit.  hist (nullable) receives max_iter * X*Y*nz values:
 * deconvRL.m:90-92 stores Xguess_history(i,:,:,:); here iteration i is the
 * contiguous block hist + i*X*Y*nz.  vdims = {X, Y, Nz}. */
int sid_deconv_rl(sid_psf *psf, const float *Htf, float *X, const int64_t vdims[3], int max_iter,
                  float *hist);
int sid_deconv_rl_device(sid_psf *psf, const float *Htf_dev, float *X_dev, const int64_t vdims[3],
                         int max_iter, float *hist_dev, void *stream);

/* --- Poisson likelihood and gradients (no reference counterpart) ---------
 * dims = {X, Y, Nz, nb}: V holds nb volumes [X,Y,Nz] back to back, y nb images
 * [X,Y] (photon counts).  nll receives sum_f sum_p F_f - y_f log F_f (FP64,
 * y log F taken as 0 where y = 0), gV (nullable) nb volumes of dNLL/dV,
 * gH (nullable) dNLL/dH in the layout of H.  All device pointers; FP32
 * projector precision only. */
int sid_poisson_nll_grad_device(sid_psf *psf, const float *V_dev, const float *y_dev, const int64_t dims[4],
                                double *nll, float *gV_dev, float *gH_dev, void *stream);
/* The FP64 normalisation of the Richardson-Lucy / EM update,
 * B1 = backwardProjectGPU_new(H, ones(X, Y)) (backwardProjectACC_new with the
 * handle's CAindex): at voxel q of plane c the exact FP64 sum of its kernel over
 * the part that lands on the sensor (summed-area tables of every kernel, built
 * plane by plane).  B1: device, column-major X*Y*Nz doubles.  No reference
 * counterpart (the reference's deconvRL has no normalisation). */
int sid_backward_ones_device(sid_psf *psf, int64_t X, int64_t Y, double *B1_dev, void *stream);
/* Vector-Jacobian product of O_f = forward(H, V_f) for upstream image
 * gradients G (nb images [X,Y]): gV = backward(G) (nullable), gH = dNLL/dH of
 * sum_f <G_f, O_f> (nullable; needs V). */
int sid_project_vjp_device(sid_psf *psf, const float *V_dev, const float *G_dev, const int64_t dims[4],
                           float *gV_dev, float *gH_dev, void *stream);

/* --- Batched library projection (src/generate_LFM_library_GPU.m:63-108) ---
 * Seed s is a sub-volume box {x0, x1, y0, y1, z0, z1} (1-based inclusive
 * sensor pixels and planes of H; boxes is n x 6 row-major), its values a
 * column-major (x1-x0+1) x (y1-y0+1) x (z1-z0+1) array.  Each seed is
 * forward-projected with planes z0..z1 of H (generate_LFM_library_GPU.m:87-88,95)
 * and its sensor image is returned on the rectangle that
 * src/project_forward_patched.m:8-20,54-59 pastes: sid_library_rects gives
 * {sx0, sx1, sy0, sy1} (1-based inclusive, n x 4) and the element offsets of
 * each column-major patch in the concatenated output (n + 1). */
int sid_library_rects(const sid_psf *psf, int64_t n_seeds, const int64_t *boxes, const int64_t sensor[2],
                      int64_t *rects, int64_t *offsets);
/* The same rectangle rule for one box without a handle (host only, no device
 * needed): n = Nnum, M = kernel side, sensor = {X, Y}. */
int sid_patched_rect(int64_t n, int64_t M, const int64_t sensor[2], const int64_t box[6], int64_t rect[4]);
/* vals_dev: the seeds' values back to back (val_offsets: n + 1 host offsets);
 * out_dev: concatenated patches.  flags bit 0: capture the launch sequence
 * in a hipGraph, cached on the handle and replayed while the seeds and the
 * buffer addresses stay the same.  One persistent kernel does the whole batch. */
int sid_library_forward_device(sid_psf *psf, int64_t n_seeds, const int64_t *boxes, const float *vals_dev,
                               const int64_t *val_offsets, const int64_t sensor[2], float *out_dev, int flags,
                               void *stream);
/* The same on host buffers (vals and out in host memory; the MEX gateway of
 * generate_LFM_library_GPU), one launch without graph capture. */
int sid_library_forward(sid_psf *psf, int64_t n_seeds, const int64_t *boxes, const float *vals,
                        const int64_t *val_offsets, const int64_t sensor[2], float *out);
/* {rows, work items, MFMA multiply-adds issued (chunks of all-zero weights are
 * skipped), accumulator elements} of the cached library plan. */
int sid_library_stats(const sid_psf *psf, int64_t stats[4]);
/* The kernel option of generate_LFM_library_GPU (src/generate_LFM_library_GPU.m:
 * 54-56,95-101: forwardFUN = forwardFUN_raw(convn(Volume, kern, 'same')) on
 * each seed's image_size x (z1-z0+1) volume).  The seed volume is zero outside
 * its box, so its convolution lives on the box grown by the kernel's reach,
 * x in [x0 - floor(kx/2), x1 + kx - 1 - floor(kx/2)] and likewise y, clipped to
 * the sensor, on the box's own planes.  sid_library_kernel_boxes gives the
 * grown boxes (n x 6, host only) and their value offsets (n + 1);
 * sid_library_convn_seeds(_device) convolves every seed's box values (FP32, the
 * library's arithmetic) into values on its grown box, which then go through
 * sid_library_forward(_device) with the grown boxes.  kdims = {kx, ky, kz} up to
 * 65 x 65 x 255, k column-major; val_offsets as for sid_library_forward. */
int sid_library_kernel_boxes(int64_t n_seeds, const int64_t *boxes, const int64_t sensor[2], const int64_t kdims[3],
                             int64_t *boxes_out, int64_t *val_offsets_out);
int sid_library_convn_seeds_device(int64_t n_seeds, const int64_t *boxes, const float *vals_dev,
                                   const int64_t *val_offsets, const int64_t sensor[2], const float *k_dev,
                                   const int64_t kdims[3], float *vals_out_dev, void *stream);
int sid_library_convn_seeds(int64_t n_seeds, const int64_t *boxes, const float *vals, const int64_t *val_offsets,
                            const int64_t sensor[2], const float *k, const int64_t kdims[3], float *vals_out);

/* --- Library of normalised footprints (src/generate_LFM_library_CPU.m:1-70) ---
 * forward_model = generate_LFM_library_CPU(centers, psf_ballistic, r, dim, m, kernel):
 * centers: n_seeds x 3 column-major doubles (MATLAB centers, 1-based x, y, z);
 * kernel == NULL: the ball of whole radius r with axial radius r / dim[2]
 * (:18,34-35); kernel [k, k, kz] column-major (kdims): its nonzero entries and
 * values (:29-32).  m = {m(1), m(2)} (the volume's lateral size).  Each centre's
 * voxels inside the volume (:43-50) are projected with the handle's PSF as
 * project_forward_patched does (:51) and divided by their L2 norm (:52-53).
 * Output: the rows of sparse(I, J, S, n_seeds, m(1)*m(2)) (:59-70) in
 * compressed-row form: row_ptr (n_seeds + 1), cols (0-based column-major pixel
 * indices, ascending within a row) and vals (FP64), at most cap entries
 * (sid_library_nsf_bound gives a sufficient cap).  A centre with no voxel inside
 * the volume, or a zero projection, gives an empty row (the reference fails in
 * project_forward_patched, or returns NaN entries).  The handle must be made
 * without CAindex and have odd Nnum and PSF side. */
int sid_library_nsf_bound(const sid_psf *psf, int64_t n_seeds, const double *centers, double r, const double dim[3],
                          const int64_t m[2], const float *kernel, const int64_t kdims[3], int64_t *bound);
int sid_library_nsf(sid_psf *psf, int64_t n_seeds, const double *centers, double r, const double dim[3],
                    const int64_t m[2], const float *kernel, const int64_t kdims[3], int64_t *row_ptr, int64_t *cols,
                    double *vals, int64_t cap);

/* --- Neuron kernel (reconstruct_S.m:110-128,184-190; generate_kernel.m) ----
 * With a kernel k set, the handle's projectors become
 *   forwardFUN(V) = F(convn(V, k, 'same')),  backwardFUN(O) = convn(B(O), k, 'same')
 * for sid_forward*, sid_backward*, sid_deconv_rl* and sid_ls_deconv* (the
 * wrapping reconstruct_S builds around forwardProjectGPU / backwardProjectGPU_new).
 * k: column-major [kx, ky, kz] host array, 1 <= kx, ky <= 65, 1 <= kz <= 255;
 * k == NULL removes the kernel.  Gradients and the batched library reject a
 * handle with a kernel (convolve seed volumes first with sid_convn_same_device). */
int sid_psf_set_kernel(sid_psf *psf, const float *k, const int64_t kdims[3]);
/* convn(V, k, 'same') on column-major volumes (MATLAB's window: offset
 * floor(size(k)/2) per dimension, zeros outside V).  Where X + floor(kx/2) and
 * Y + floor(ky/2) are at most 2048 and kz <= 8: FP32 2-D transforms in (x, y)
 * with the z sum per frequency, at every precision (rel-L2 about 1e-7, below
 * the FP32 direct sum's rounding).  Otherwise, or with SID_CONVN_FFT=0 in the
 * environment: SID_PREC_FP32 / SID_PREC_FP64 the FP32 direct sum,
 * SID_PREC_BF16X2 / SID_PREC_BF16X3 BF16-split MFMA (volume hi + lo, kernel
 * BF16, or hi + lo for BF16X3).  A handle with a kernel does the same.  _device: V, k, out device
 * pointers on `stream`; the plain variant takes host arrays (device 0).  The
 * transform form keeps its work arrays per device for the process (calls on
 * one device take turns), and a non-finite input value makes NaN every (x, y)
 * of the output planes within the kernel's depth (the FP32 direct sum spreads
 * it over the kernel's footprint only). */
int sid_convn_same_device(const float *V_dev, const int64_t vdims[3], const float *k_dev, const int64_t kdims[3],
                          int precision, float *out_dev, void *stream);
int sid_convn_same(const float *V, const int64_t vdims[3], const float *k, const int64_t kdims[3], int precision,
                   float *out);

/* --- Movie background and std image (sid_main.m:88-103) ------------------
 * Y: column-major movie [P pixels, T frames] (one frame per column), read only.
 *   flags & 1: work on Y / max(Y(:)) (sid_main.m:88-89); *ymax receives the
 *              maximum (0 when not normalising; ymax may be NULL)
 *   bg_iter > 0: [s, t] = rank_1_factorization(Y, bg_iter)
 *              (src/rank_1_factorization.m:31-39); bg_iter == 0: s = 0, t = 0
 *              (bg_sub off, sid_main.m:94-98)
 *   std = compute_std_image(Y, s, t) (src/compute_std_image.m:22-26); a
 *              negative radicand (rounding) gives 0 where MATLAB returns a complex value
 * Outputs: s [P], t [T], std [P] (FP32; s and t may be NULL).  FP64
 * accumulation, deterministic.  _device: device pointers on `stream`; the plain
 * variant takes host arrays (device 0). */
int sid_std_image_device(const float *Y_dev, int64_t P, int64_t T, int bg_iter, int flags, float *s_dev,
                         float *t_dev, float *std_dev, double *ymax, void *stream);
int sid_std_image(const float *Y, int64_t P, int64_t T, int bg_iter, int flags, float *s, float *t, float *std_img,
                  double *ymax);
/* sid_std_image_device without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under stream capture,
 * serving loops, harnesses that forbid hipMalloc).  The same kernels and
 * results as sid_std_image_device; its work space (block partial sums, the FP64
 * s and t vectors, per-pixel sums: about 24 B per pixel + 8 B per frame and
 * pixel chunk) lives in a 32 MiB array the library holds on every device it
 * runs on until the process ends (enough for 1005^2 px x 1000 frames).  Larger
 * movies are rejected with SID_ERR_DIMS.  Every call on one device shares that
 * array: calls that overlap on one device race. */
int sid_std_image_device_noalloc(const float *Y_dev, int64_t P, int64_t T, int bg_iter, int flags, float *s_dev,
                                 float *t_dev, float *std_dev, double *ymax, void *stream);

/* --- LS_deconv (external/LFrecon_util/LS_deconv.m:1-97) ------------------
 * Projected gradient descent with exact line search for
 *   argmin_{x >= 0} ||y - A x||^2 + TV_L2 + lamb_L2 ||x||^2 + lamb_L1 ||x||_1,
 * started at Xguess = df (= A'y), with Q(x) = A'A x + tv(x) + lamb_L2 x where
 * tv is the 7-point 'same' stencil of LS_deconv.m:54-58 (on when
 * max(lamb_TV_L2) != 0).  The step is d'd / ||A d||^2, or d'd / d'Q(d) when
 * lamb_TV_L1 * lamb_TV_L2 * lamb_L2 is nonzero in all three entries (:80-84).
 * Dot products are FP64.  opts == NULL means the reference defaults
 * (max_iter 8, all multipliers 0).  X receives the result; vdims = {X, Y, Nz}. */
typedef struct sid_ls_opts {
    int max_iter;          /* opts.max_iter */
    double lamb_TV_L1;     /* opts.lamb_TV_L1 (only selects the step formula) */
    double lamb_TV_L2[3];  /* opts.lamb_TV_L2, per dimension of the volume */
    double lamb_L1;        /* opts.lamb_L1 */
    double lamb_L2;        /* opts.lamb_L2 */
} sid_ls_opts;
int sid_ls_deconv(sid_psf *psf, const float *df, float *X, const int64_t vdims[3], const sid_ls_opts *opts);
int sid_ls_deconv_device(sid_psf *psf, const float *df_dev, float *X_dev, const int64_t vdims[3],
                         const sid_ls_opts *opts, void *stream);

/* Device time (ms, HIP events) of the last call's projector kernels:
 * [0] forward direct GEMM, [1] forward reduce, [2] backward direct GEMM,
 * [3] backward reduce + update, [4] whole call, [5] forward FFT path,
 * [6] backward FFT path, [7] forward spectra GEMV, [8] backward spectra GEMV,
 * [9..11] 0.  Summed over iterations for sid_deconv_rl.  sid_ls_deconv and
 * the ISRA loop of a handle with a neuron kernel queue their projections
 * without per-projection timing and leave these values as they were. */
int sid_last_timing(const sid_psf *psf, double ms[12]);

/* Kernel launch counts and GEMM geometry of the direct path for the current
 * sensor plan (for roofline accounting): {macs_fwd, macs_bwd, items_fwd,
 * items_bwd, slabs_fwd, slabs_bwd, rows_padded, cols_padded}.  macs = the MFMA
 * multiply-adds one launch runs (K rows executed after the all-zero chunks are
 * skipped, x padded N x padded M; one product per MAC, so BF16X2 issues twice
 * as many MFMA operations and BF16X3 three times). */
int sid_plan_stats(const sid_psf *psf, int64_t stats[8]);
/* The same MFMA multiply-adds per plane of H (n = size(H,5); 0 for planes on
 * the FFT path or empty) for the forward and backward launches. */
int sid_plan_plane_macs(const sid_psf *psf, int64_t *fwd, int64_t *bwd, int64_t n);

/* The FFT primitive of the FFT path: nseq contiguous complex sequences
 * (interleaved re, im floats) of length L (2/3/5-smooth, <= 2048) transformed
 * in place on the device; direction -1 forward, +1 inverse (unnormalised). */
int sid_fft_device(int L, int64_t nseq, float *data_dev, int direction, void *stream);

/* Built-in check of the MFMA lane layouts the kernels assume (one 16x16x32
 * BF16 and one 16x16x4 FP32 product against a host loop).  0 = pass. */
int sid_selftest(void);

#ifdef __cplusplus
}
#endif
#endif /* SID_PROJ_H */
