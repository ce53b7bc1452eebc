/*
 * MEX gateways over libsidproj (include/sid_proj.h) with the reference's
 * MATLAB names and argument order.  One source, built once per entry point:
 *
 *   mex -DSID_MEX_FORWARD  -output forwardProjectGPU       sid_mex_gateway.c -lsidproj
 *   mex -DSID_MEX_BACKWARD -output backwardProjectGPU_new  sid_mex_gateway.c -lsidproj
 *   mex -DSID_MEX_FORWARD  -DSID_MEX_CROP -output forwardProjectACC      sid_mex_gateway.c -lsidproj
 *   mex -DSID_MEX_BACKWARD -DSID_MEX_CROP -output backwardProjectACC_new sid_mex_gateway.c -lsidproj
 *   mex -DSID_MEX_ISRA     -output sidDeconvRL             sid_mex_gateway.c -lsidproj
 *   mex -DSID_MEX_LS       -output sidLSDeconv             sid_mex_gateway.c -lsidproj
 *   mex -DSID_MEX_LIBGPU   -output generate_LFM_library_GPU sid_mex_gateway.c -lsidproj
 *   mex -DSID_MEX_LIBCPU   -output generate_LFM_library_CPU sid_mex_gateway.c -lsidproj
 *   mex -DSID_MEX_PSF      -output sidPSF                  sid_mex_gateway.c -lsidproj
 *     (h = sidPSF('create', H [, CAindex]); 'precision', 'kernel' and 'destroy' act on h)
 *
 * Replaces: external/LFrecon_util/forwardProjectGPU.m:1 (O = forwardProjectGPU(H, realspace)),
 * backwardProjectGPU_new.m:1 (B = backwardProjectGPU_new(H, projection)), forwardProjectACC.m:1 and
 * backwardProjectACC_new.m:1 (third argument CAindex; the ACC backward returns double,
 * backwardProjectACC_new.m:6), the loops of deconvRL.m:1-93 ([X, Xguess_history] =
 * sidDeconvRL(H, Htf, maxIter, Xguess) = deconvRL(@(x) forwardProjectGPU(H,x),
 * @(y) backwardProjectGPU_new(H,y), Htf, maxIter, Xguess), opts.mode 'basic', lambda 0) and of
 * LS_deconv.m:1-97 (X = sidLSDeconv(H, df, opts)), src/generate_LFM_library_GPU.m:1 and
 * src/generate_LFM_library_CPU.m:1 (same arguments and outputs).
 *
 * H is an array [M, M, Nnum, Nnum, Nz], a struct with field H (psf_ballistic), or a
 * handle h = sidPSF('create', H [, CAindex]) (uint64; sidPSF('destroy', h) releases it,
 * sidPSF('precision', h, 'fp32' | 'bf16x2' | 'bf16x3' | 'fp64') selects the arithmetic).
 * Arrays are MATLAB column-major, exactly the layout the C ABI takes; double inputs are
 * converted to single (forwardProjectGPU.m:6,21 computes in single).  For an array H the
 * gateway keeps one device handle, reused while a 64-bit hash of all of H's values, its
 * dimensions and the CAindex crop are unchanged (hashing reads H once per call; a handle
 * argument skips that); it is released at mexAtExit.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"
#include "sid_proj.h"

static void fail(int rc, const char *where)
{
    char id[64];
    static const char *names[] = {"ok", "arg", "dims", "nz", "oom", "hip", "state", "nodev"};
    snprintf(id, sizeof id, "sid:%s", (rc >= 0 && rc < 8) ? names[rc] : "unknown");
    mexErrMsgIdAndTxt(id, "%s: %s", where, sid_last_error());
}

static void arg_error(const char *msg) { mexErrMsgIdAndTxt("sid:arg", "%s", msg); }

static int is_real_numeric(const mxArray *a)
{
    return (mxIsSingle(a) || mxIsDouble(a)) && !mxIsComplex(a) && !mxIsSparse(a);
}

/* single-precision copy of a real single or double array (mxMalloc'ed when converted) */
static float *as_single(const mxArray *a, int *owned)
{
    const mwSize n = mxGetNumberOfElements(a);
    if (!is_real_numeric(a)) arg_error("expected a real single or double array");
    if (mxIsSingle(a)) {
        *owned = 0;
        return (float *)mxGetData(a);
    }
    {
        const double *d = (const double *)mxGetData(a);
        float *f = (float *)mxMalloc((n ? n : 1) * sizeof(float));
        mwSize i;
        for (i = 0; i < n; ++i) f[i] = (float)d[i];
        *owned = 1;
        return f;
    }
}

/* double-precision copy of a real single or double array (always mxMalloc'ed) */
static double *as_double(const mxArray *a)
{
    const mwSize n = mxGetNumberOfElements(a);
    double *d = (double *)mxMalloc((n ? n : 1) * sizeof(double));
    mwSize i;
    if (!is_real_numeric(a)) arg_error("expected a real single or double array");
    if (mxIsDouble(a)) memcpy(d, mxGetData(a), n * sizeof(double));
    else
        for (i = 0; i < n; ++i) d[i] = ((const float *)mxGetData(a))[i];
    return d;
}

static double scalar_of(const mxArray *a, const char *what)
{
    char msg[128];
    if (!is_real_numeric(a) || mxGetNumberOfElements(a) != 1) {
        snprintf(msg, sizeof msg, "%s must be a real scalar", what);
        arg_error(msg);
    }
    return mxGetScalar(a);
}

static void dims_of(const mxArray *a, int64_t *out, int k)
{
    const mwSize nd = mxGetNumberOfDimensions(a);
    const mwSize *d = mxGetDimensions(a);
    int i;
    for (i = 0; i < k; ++i) out[i] = (i < (int)nd) ? (int64_t)d[i] : 1;
    for (i = k; i < (int)nd; ++i) out[k - 1] *= (int64_t)d[i];   /* trailing dims fold into the last */
}

/* 64-bit hash of n bytes: four lanes of multiply / rotate over 8-byte words, then
 * the tail bytes; any change of a value changes its lane with probability ~1 - 2^-64. */
static uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
static uint64_t hash_bytes(const void *p, size_t n, uint64_t h)
{
    const unsigned char *b = (const unsigned char *)p;
    const uint64_t K1 = 0x9E3779B185EBCA87ull, K2 = 0xC2B2AE3D27D4EB4Full;
    uint64_t v[4] = {h ^ K1, h ^ K2, h + K1, h - K2};
    size_t i = 0;
    for (; i + 32 <= n; i += 32) {
        int l;
        for (l = 0; l < 4; ++l) {
            uint64_t w;
            memcpy(&w, b + i + 8 * l, 8);
            v[l] = rotl64(v[l] + w * K2, 31) * K1;
        }
    }
    h = rotl64(v[0], 1) + rotl64(v[1], 7) + rotl64(v[2], 12) + rotl64(v[3], 18) + (uint64_t)n;
    for (; i < n; ++i) h = rotl64(h ^ (b[i] * K1), 11) * K2;
    h ^= h >> 33;
    h *= K2;
    h ^= h >> 29;
    return h;
}

/* ---- PSF handles --------------------------------------------------------- */

static sid_psf *g_psf = NULL;   /* the gateway's own handle for an array H */
static uint64_t g_key = 0;

static void release(void)
{
    if (g_psf) sid_psf_destroy(g_psf);
    g_psf = NULL;
}

static int32_t *caindex_of(const mxArray *CAa, int64_t nz)
{
    const mwSize m = mxGetNumberOfElements(CAa);
    int32_t *CA;
    double *d;
    mwSize i;
    if (!is_real_numeric(CAa) || (int64_t)m != 2 * nz) arg_error("CAindex must be a real Nz x 2 array");
    d = as_double(CAa);
    CA = (int32_t *)mxMalloc(m * sizeof(int32_t));
    for (i = 0; i < m; ++i) CA[i] = (int32_t)d[i];
    mxFree(d);
    return CA;
}

static const mxArray *psf_array(const mxArray *Ha)
{
    if (mxIsStruct(Ha)) {
        const mxArray *f = mxGetField(Ha, 0, "H");
        if (!f) arg_error("the PSF struct has no field H");
        return f;
    }
    return Ha;
}

/* A handle argument (uint64 scalar from sidPSF('create', ...)), checked live. */
static sid_psf *handle_arg(const mxArray *a)
{
    sid_psf *p;
    uint64_t v;
    if (!mxIsUint64(a) || mxGetNumberOfElements(a) != 1) arg_error("expected a sidPSF handle (uint64 scalar)");
    memcpy(&v, mxGetData(a), sizeof v);
    p = (sid_psf *)(uintptr_t)v;
    if (sid_psf_is_live(p) != SID_OK) arg_error("not a live sidPSF handle");
    return p;
}

static sid_psf *psf_for(const mxArray *Ha, const mxArray *CAa)
{
    int64_t dims[5];
    int owned, rc;
    float *H;
    int32_t *CA = NULL;
    uint64_t key;
    if (mxIsUint64(Ha)) {
        if (CAa) arg_error("pass CAindex to sidPSF('create', H, CAindex), not with a handle");
        return handle_arg(Ha);
    }
    Ha = psf_array(Ha);
    if (mxGetNumberOfDimensions(Ha) > 5) arg_error("H must be [M, M, Nnum, Nnum, Nz]");
    dims_of(Ha, dims, 5);
    H = as_single(Ha, &owned);
    key = hash_bytes(dims, sizeof dims, 0x51D);
    key = hash_bytes(H, (size_t)mxGetNumberOfElements(Ha) * sizeof(float), key);
    if (CAa) {
        CA = caindex_of(CAa, dims[4]);
        key = hash_bytes(CA, (size_t)(2 * dims[4]) * sizeof(int32_t), key ^ 1);
    }
    if (g_psf && g_key == key) {
        if (owned) mxFree(H);
        if (CA) mxFree(CA);
        return g_psf;
    }
    release();
    rc = sid_psf_create(H, dims, CA, &g_psf);
    if (owned) mxFree(H);
    if (CA) mxFree(CA);
    if (rc != SID_OK) {
        g_psf = NULL;
        fail(rc, "sid_psf_create");
    }
    g_key = key;
    mexAtExit(release);
    return g_psf;
}

#if defined(SID_MEX_LS)
/* opts.<name> as a double (default when absent) */
static double field_or(const mxArray *opts, const char *name, double dflt)
{
    const mxArray *f = opts ? mxGetField(opts, 0, name) : NULL;
    return f ? scalar_of(f, name) : dflt;
}
#endif

#if defined(SID_MEX_LIBGPU) || defined(SID_MEX_LIBCPU)
static double round_half_away(double x) { return x < 0 ? -(double)(int64_t)(-x + 0.5) : (double)(int64_t)(x + 0.5); }
#endif

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
#if defined(SID_MEX_FORWARD)
    {
        /* O = forwardProjectGPU(H, realspace)  /  forwardProjectACC(H, realspace, CAindex) */
        int64_t vd[3];
        int owned, rc;
        float *V;
        mwSize od[2];
        sid_psf *p;
        (void)nlhs;
#ifdef SID_MEX_CROP
        if (nrhs != 3) arg_error("usage: O = forwardProjectACC(H, realspace, CAindex)");
        p = psf_for(prhs[0], prhs[2]);
#else
        if (nrhs != 2) arg_error("usage: O = forwardProjectGPU(H, realspace)");
        p = psf_for(prhs[0], NULL);
#endif
        if (mxGetNumberOfDimensions(prhs[1]) > 3) arg_error("realspace must be [X, Y, nz]");
        dims_of(prhs[1], vd, 3);
        V = as_single(prhs[1], &owned);
        od[0] = (mwSize)vd[0];
        od[1] = (mwSize)vd[1];
        plhs[0] = mxCreateNumericArray(2, od, mxSINGLE_CLASS, mxREAL);
        rc = sid_forward(p, V, vd, (float *)mxGetData(plhs[0]));
        if (owned) mxFree(V);
        if (rc != SID_OK) fail(rc, "forwardProjectGPU");
    }
#elif defined(SID_MEX_BACKWARD)
    {
        /* B = backwardProjectGPU_new(H, projection) (single)  /  backwardProjectACC_new(H, projection, CAindex)
         * (double: backwardProjectACC_new.m:6 allocates zeros(...) of class double) */
        int64_t odims[2], nzH;
        int owned, rc;
        float *O;
        mwSize bd[3];
        sid_psf *p;
        sid_psf_info *info;
        (void)nlhs;
#ifdef SID_MEX_CROP
        if (nrhs != 3) arg_error("usage: B = backwardProjectACC_new(H, projection, CAindex)");
        p = psf_for(prhs[0], prhs[2]);
#else
        if (nrhs != 2) arg_error("usage: B = backwardProjectGPU_new(H, projection)");
        p = psf_for(prhs[0], NULL);
#endif
        info = (sid_psf_info *)mxMalloc(sizeof(sid_psf_info));
        rc = sid_psf_get_info(p, info);
        nzH = info->Nz;
        mxFree(info);
        if (rc != SID_OK) fail(rc, "sid_psf_get_info");
        if (mxGetNumberOfDimensions(prhs[1]) > 2) arg_error("projection must be [X, Y]");
        dims_of(prhs[1], odims, 2);
        O = as_single(prhs[1], &owned);
        bd[0] = (mwSize)odims[0];
        bd[1] = (mwSize)odims[1];
        bd[2] = (mwSize)nzH;
#ifdef SID_MEX_CROP
        {
            const size_t nb = (size_t)(odims[0] * odims[1] * nzH);
            float *tmp = (float *)mxMalloc((nb ? nb : 1) * sizeof(float));
            size_t i;
            double *out;
            rc = sid_backward(p, O, odims, tmp);
            if (owned) mxFree(O);
            if (rc != SID_OK) {
                mxFree(tmp);
                fail(rc, "backwardProjectACC_new");
            }
            plhs[0] = mxCreateNumericArray(3, bd, mxDOUBLE_CLASS, mxREAL);
            out = (double *)mxGetData(plhs[0]);
            for (i = 0; i < nb; ++i) out[i] = tmp[i];
            mxFree(tmp);
        }
#else
        plhs[0] = mxCreateNumericArray(3, bd, mxSINGLE_CLASS, mxREAL);
        rc = sid_backward(p, O, odims, (float *)mxGetData(plhs[0]));
        if (owned) mxFree(O);
        if (rc != SID_OK) fail(rc, "backwardProjectGPU_new");
#endif
    }
#elif defined(SID_MEX_ISRA)
    {
        /* [X, Xguess_history] = sidDeconvRL(H, Htf, maxIter, Xguess) */
        int64_t vd[3], nv;
        int oh, ox, rc, it;
        float *Htf, *X0, *hist = NULL;
        mwSize xd[3];
        sid_psf *p;
        if (nrhs != 4) arg_error("usage: [X, Xguess_history] = sidDeconvRL(H, Htf, maxIter, Xguess)");
        if (mxGetNumberOfDimensions(prhs[1]) > 3) arg_error("Htf must be [X, Y, Nz]");
        dims_of(prhs[1], vd, 3);
        nv = vd[0] * vd[1] * vd[2];
        if ((int64_t)mxGetNumberOfElements(prhs[3]) != nv) arg_error("Xguess must have the size of Htf");
        {
            const double mi = scalar_of(prhs[2], "maxIter");
            if (!(mi >= 0) || mi != (double)(int)mi) arg_error("maxIter must be a whole number >= 0");
            it = (int)mi;
        }
        p = psf_for(prhs[0], NULL);
        Htf = as_single(prhs[1], &oh);
        X0 = as_single(prhs[3], &ox);
        xd[0] = (mwSize)vd[0];
        xd[1] = (mwSize)vd[1];
        xd[2] = (mwSize)vd[2];
        plhs[0] = mxCreateNumericArray(3, xd, mxSINGLE_CLASS, mxREAL);
        memcpy(mxGetData(plhs[0]), X0, (size_t)nv * sizeof(float));
        if (nlhs > 1 && it > 0) hist = (float *)mxMalloc((size_t)nv * it * sizeof(float));
        rc = sid_deconv_rl(p, Htf, (float *)mxGetData(plhs[0]), vd, it, hist);
        if (oh) mxFree(Htf);
        if (ox) mxFree(X0);
        if (rc != SID_OK) {
            if (hist) mxFree(hist);
            fail(rc, "sidDeconvRL");
        }
        if (nlhs > 1) {
            /* Xguess_history = zeros([maxIter size(Xguess)], 'single'); (i, :, :, :) = iterate i
             * (deconvRL.m:9-11,90-92); the library returns iterate i as one contiguous block */
            mwSize hd[4];
            float *out;
            int64_t i, e;
            hd[0] = (mwSize)it;
            hd[1] = xd[0];
            hd[2] = xd[1];
            hd[3] = xd[2];
            plhs[1] = mxCreateNumericArray(4, hd, mxSINGLE_CLASS, mxREAL);
            out = (float *)mxGetData(plhs[1]);
            for (i = 0; i < it; ++i)
                for (e = 0; e < nv; ++e) out[i + (int64_t)it * e] = hist[i * nv + e];
            if (hist) mxFree(hist);
        }
    }
#elif defined(SID_MEX_LS)
    {
        /* X = sidLSDeconv(H, df, opts) = LS_deconv(@(x) forwardProjectGPU(H,x), @(y) backwardProjectGPU_new(H,y),
         * df, opts); opts fields as LS_deconv.m:26-43 (lamb_TV_L2 may be a scalar, used for all three axes) */
        int64_t vd[3], nv;
        int od, rc, k;
        float *df;
        mwSize xd[3];
        sid_psf *p;
        sid_ls_opts o;
        const mxArray *opts = nrhs > 2 ? prhs[2] : NULL;
        (void)nlhs;
        if (nrhs < 2 || nrhs > 3) arg_error("usage: X = sidLSDeconv(H, df, opts)");
        if (opts && !mxIsStruct(opts)) arg_error("opts must be a struct");
        o.max_iter = (int)field_or(opts, "max_iter", 8);
        o.lamb_TV_L1 = field_or(opts, "lamb_TV_L1", 0);
        o.lamb_L1 = field_or(opts, "lamb_L1", 0);
        o.lamb_L2 = field_or(opts, "lamb_L2", 0);
        o.lamb_TV_L2[0] = o.lamb_TV_L2[1] = o.lamb_TV_L2[2] = 0;
        if (opts && mxGetField(opts, 0, "lamb_TV_L2")) {
            const mxArray *t = mxGetField(opts, 0, "lamb_TV_L2");
            const mwSize nt = mxGetNumberOfElements(t);
            double *tv;
            if (!is_real_numeric(t) || (nt != 1 && nt != 3)) arg_error("opts.lamb_TV_L2 must have 1 or 3 entries");
            tv = as_double(t);
            for (k = 0; k < 3; ++k) o.lamb_TV_L2[k] = tv[nt == 1 ? 0 : k];
            mxFree(tv);
        }
        if (mxGetNumberOfDimensions(prhs[1]) > 3) arg_error("df must be [X, Y, Nz]");
        dims_of(prhs[1], vd, 3);
        nv = vd[0] * vd[1] * vd[2];
        p = psf_for(prhs[0], NULL);
        df = as_single(prhs[1], &od);
        xd[0] = (mwSize)vd[0];
        xd[1] = (mwSize)vd[1];
        xd[2] = (mwSize)vd[2];
        plhs[0] = mxCreateNumericArray(3, xd, mxSINGLE_CLASS, mxREAL);
        (void)nv;
        rc = sid_ls_deconv(p, df, (float *)mxGetData(plhs[0]), vd, &o);
        if (od) mxFree(df);
        if (rc != SID_OK) fail(rc, "sidLSDeconv");
    }
#elif defined(SID_MEX_LIBGPU)
    {
        /* forward_model = generate_LFM_library_GPU(recon, centers, neur_id, psf, options)
         * (src/generate_LFM_library_GPU.m:1-113): per centre, the mean of the listed reconstructions on the
         * clipped box (:68-85), projected with planes z0..z1 of psf.H (:87-95); row k of the single
         * N x prod(options.image_size) result is that projection (:104-106), here from one batched launch */
        const mxArray *recon = prhs[0], *opts;
        int64_t N, K, X, Y, Nz, r, dz, s, k;
        double *cen, *ids, axial;
        int64_t *boxes, *voff, *rects, *offs, sensor[2], vdims[3];
        float *vals, *patches;
        sid_psf *p;
        int rc;
        mwSize fd[2];
        float *fm;
        (void)nlhs;
        if (nrhs != 5) arg_error("usage: forward_model = generate_LFM_library_GPU(recon, centers, neur_id, psf, options)");
        if (!mxIsCell(recon) || mxGetNumberOfElements(recon) < 1) arg_error("recon must be a non-empty cell array");
        opts = prhs[4];
        if (!mxIsStruct(opts)) arg_error("options must be a struct");
        if (!mxGetField(opts, 0, "image_size") || !mxGetField(opts, 0, "axial") || !mxGetField(opts, 0, "neur_rad"))
            arg_error("options needs image_size, axial and neur_rad");
        {
            double *is = as_double(mxGetField(opts, 0, "image_size"));
            if (mxGetNumberOfElements(mxGetField(opts, 0, "image_size")) < 2) arg_error("options.image_size needs 2 entries");
            X = (int64_t)is[0];
            Y = (int64_t)is[1];
            mxFree(is);
        }
        axial = scalar_of(mxGetField(opts, 0, "axial"), "options.axial");
        r = (int64_t)scalar_of(mxGetField(opts, 0, "neur_rad"), "options.neur_rad");
        dz = (int64_t)ceil((double)r / axial);
        if (mxGetNumberOfDimensions(prhs[1]) != 2 || mxGetN(prhs[1]) != 3) arg_error("centers must be N x 3");
        N = (int64_t)mxGetM(prhs[1]);
        K = (int64_t)mxGetNumberOfElements(recon);
        if ((int64_t)mxGetM(prhs[2]) != N || (int64_t)mxGetN(prhs[2]) > K)
            arg_error("neur_id must have one row per centre and at most numel(recon) columns");
        p = psf_for(prhs[3], NULL);
        {
            sid_psf_info *info = (sid_psf_info *)mxMalloc(sizeof(sid_psf_info));
            rc = sid_psf_get_info(p, info);
            Nz = info->Nz;
            mxFree(info);
            if (rc != SID_OK) fail(rc, "sid_psf_get_info");
        }
        cen = as_double(prhs[1]);
        ids = as_double(prhs[2]);
        for (k = 0; k < K; ++k) {
            const mxArray *c = mxGetCell(recon, (mwIndex)k);
            if (!c || !is_real_numeric(c)) arg_error("recon{k} must be real single or double arrays");
            dims_of(c, vdims, 3);
            if (vdims[0] != X || vdims[1] != Y || vdims[2] != Nz)
                arg_error("recon{k} must be image_size(1) x image_size(2) x size(psf.H,5)");
        }
        boxes = (int64_t *)mxMalloc(6 * (N ? N : 1) * sizeof(int64_t));
        voff = (int64_t *)mxMalloc((N + 1) * sizeof(int64_t));
        voff[0] = 0;
        for (s = 0; s < N; ++s) {
            const int64_t x = (int64_t)round_half_away(cen[s]), y = (int64_t)round_half_away(cen[s + N]),
                          z = (int64_t)round_half_away(cen[s + 2 * N]);
            int64_t *b = boxes + 6 * s;
            b[0] = x - r > 1 ? x - r : 1;
            b[1] = x + r < X ? x + r : X;
            b[2] = y - r > 1 ? y - r : 1;
            b[3] = y + r < Y ? y + r : Y;
            b[4] = z - dz > 1 ? z - dz : 1;
            b[5] = z + dz < Nz ? z + dz : Nz;
            if (b[1] < b[0] || b[3] < b[2] || b[5] < b[4]) arg_error("a centre's box lies outside the volume");
            voff[s + 1] = voff[s] + (b[1] - b[0] + 1) * (b[3] - b[2] + 1) * (b[5] - b[4] + 1);
        }
        vals = (float *)mxMalloc((voff[N] ? voff[N] : 1) * sizeof(float));
        for (s = 0; s < N; ++s) {
            const int64_t *b = boxes + 6 * s;
            const int64_t wx = b[1] - b[0] + 1, wy = b[3] - b[2] + 1, wz = b[5] - b[4] + 1;
            int64_t nid = 0, e, i, j, l;
            double *acc = (double *)mxMalloc((size_t)(wx * wy * wz) * sizeof(double));
            for (e = 0; e < wx * wy * wz; ++e) acc[e] = 0.0;
            for (k = 0; k < (int64_t)mxGetN(prhs[2]); ++k) {
                const mxArray *c;
                if (ids[s + N * k] == 0.0) continue;   /* ids = find(neur_id(kk, :)) (:74-75) */
                ++nid;
                c = mxGetCell(recon, (mwIndex)k);
                for (l = 0; l < wz; ++l)
                    for (j = 0; j < wy; ++j)
                        for (i = 0; i < wx; ++i) {
                            const int64_t src = (b[0] - 1 + i) + X * ((b[2] - 1 + j) + Y * (b[4] - 1 + l));
                            const double v = mxIsSingle(c) ? ((const float *)mxGetData(c))[src]
                                                           : ((const double *)mxGetData(c))[src];
                            acc[i + wx * (j + wy * l)] += v;
                        }
            }
            for (e = 0; e < wx * wy * wz; ++e) vals[voff[s] + e] = (float)(acc[e] / (double)nid);   /* :85 */
            mxFree(acc);
        }
        mxFree(cen);
        mxFree(ids);
        sensor[0] = X;
        sensor[1] = Y;
        /* kernel = generate_kernel(options.ker_shape, options.ker_param) when ker_param is set (:50-56; the
         * reference's own generate_kernel.m, called back), forwardFUN = forwardFUN_raw(convn(V, kern,
         * 'same')) (:95-101): each seed moves to its box grown by the kernel's reach, with the convolved
         * values */
        if (mxGetField(opts, 0, "ker_param") && mxGetNumberOfElements(mxGetField(opts, 0, "ker_param")) > 0) {
            mxArray *kin[2], *kout[1] = {NULL};
            mxArray *ks = mxGetField(opts, 0, "ker_shape");
            int64_t kd[3], *gboxes, *goff;
            float *kern, *gvals;
            int kown = 0;
            kin[0] = ks ? ks : mxCreateString("ball");
            kin[1] = mxGetField(opts, 0, "ker_param");
            mexCallMATLAB(1, kout, 2, kin, "generate_kernel");
            if (!kout[0] || !is_real_numeric(kout[0]) || mxGetNumberOfDimensions(kout[0]) > 3)
                arg_error("generate_kernel must return a real [kx, ky, kz] array");
            dims_of(kout[0], kd, 3);
            kern = as_single(kout[0], &kown);
            gboxes = (int64_t *)mxMalloc(6 * (N ? N : 1) * sizeof(int64_t));
            goff = (int64_t *)mxMalloc((N + 1) * sizeof(int64_t));
            rc = sid_library_kernel_boxes(N, boxes, sensor, kd, gboxes, goff);
            if (rc != SID_OK) fail(rc, "sid_library_kernel_boxes");
            gvals = (float *)mxMalloc((goff[N] ? goff[N] : 1) * sizeof(float));
            rc = sid_library_convn_seeds(N, boxes, vals, voff, sensor, kern, kd, gvals);
            if (kown) mxFree(kern);
            if (rc != SID_OK) fail(rc, "sid_library_convn_seeds");
            mxFree(boxes);
            mxFree(voff);
            mxFree(vals);
            boxes = gboxes;
            voff = goff;
            vals = gvals;
        }
        rects = (int64_t *)mxMalloc(4 * (N ? N : 1) * sizeof(int64_t));
        offs = (int64_t *)mxMalloc((N + 1) * sizeof(int64_t));
        rc = sid_library_rects(p, N, boxes, sensor, rects, offs);
        if (rc != SID_OK) fail(rc, "sid_library_rects");
        patches = (float *)mxMalloc((offs[N] ? offs[N] : 1) * sizeof(float));
        rc = sid_library_forward(p, N, boxes, vals, voff, sensor, patches);
        mxFree(vals);
        if (rc != SID_OK) fail(rc, "generate_LFM_library_GPU");
        fd[0] = (mwSize)N;
        fd[1] = (mwSize)(X * Y);
        plhs[0] = mxCreateNumericArray(2, fd, mxSINGLE_CLASS, mxREAL);
        fm = (float *)mxGetData(plhs[0]);
        for (s = 0; s < N; ++s) {
            const int64_t *rc4 = rects + 4 * s, w = rc4[1] - rc4[0] + 1, h = rc4[3] - rc4[2] + 1;
            int64_t i, j;
            for (j = 0; j < h; ++j)
                for (i = 0; i < w; ++i)
                    fm[s + N * ((rc4[0] - 1 + i) + X * (rc4[2] - 1 + j))] = patches[offs[s] + i + w * j];
        }
        mxFree(patches);
        mxFree(rects);
        mxFree(offs);
        mxFree(boxes);
        mxFree(voff);
    }
#elif defined(SID_MEX_LIBCPU)
    {
        /* forward_model = generate_LFM_library_CPU(centers, psf_ballistic, r, dim, m [, kernel])
         * (src/generate_LFM_library_CPU.m:1-73): a sparse double N x prod(m(1:2)) */
        int64_t N, m2[2], bound = 0, kd[3], nnz, s, j, c;
        double *cen, *dim, r;
        const mxArray *ker = nrhs > 5 ? prhs[5] : NULL;
        float *kf = NULL;
        int kown = 0, rc;
        int64_t *row_ptr, *cols, *cnt;
        double *vals, *pr;
        mwIndex *ir, *jc;
        sid_psf *p;
        (void)nlhs;
        if (nrhs < 5 || nrhs > 6) arg_error("usage: forward_model = generate_LFM_library_CPU(centers, psf_ballistic, r, dim, m [, kernel])");
        if (mxGetNumberOfDimensions(prhs[0]) != 2 || mxGetN(prhs[0]) != 3) arg_error("centers must be N x 3");
        if (mxGetNumberOfElements(prhs[3]) < 3) arg_error("dim must have 3 entries");
        if (mxGetNumberOfElements(prhs[4]) < 2) arg_error("m must have at least 2 entries");
        N = (int64_t)mxGetM(prhs[0]);
        cen = as_double(prhs[0]);
        dim = as_double(prhs[3]);
        {
            double *mm = as_double(prhs[4]);
            m2[0] = (int64_t)mm[0];
            m2[1] = (int64_t)mm[1];
            mxFree(mm);
        }
        r = scalar_of(prhs[2], "r");
        if (ker) {
            if (mxGetNumberOfDimensions(ker) > 3) arg_error("kernel must be [k, k, kz]");
            dims_of(ker, kd, 3);
            kf = as_single(ker, &kown);
        }
        p = psf_for(prhs[1], NULL);
        rc = sid_library_nsf_bound(p, N, cen, r, dim, m2, kf, ker ? kd : NULL, &bound);
        if (rc != SID_OK) fail(rc, "generate_LFM_library_CPU");
        row_ptr = (int64_t *)mxMalloc((N + 1) * sizeof(int64_t));
        cols = (int64_t *)mxMalloc((bound ? bound : 1) * sizeof(int64_t));
        vals = (double *)mxMalloc((bound ? bound : 1) * sizeof(double));
        rc = sid_library_nsf(p, N, cen, r, dim, m2, kf, ker ? kd : NULL, row_ptr, cols, vals, bound);
        mxFree(cen);
        mxFree(dim);
        if (kown) mxFree(kf);
        if (rc != SID_OK) fail(rc, "generate_LFM_library_CPU");
        /* compressed rows (centre-major) -> MATLAB's compressed columns (pixel-major) */
        nnz = row_ptr[N];
        plhs[0] = mxCreateSparse((mwSize)N, (mwSize)(m2[0] * m2[1]), (mwSize)(nnz ? nnz : 1), mxREAL);
        pr = mxGetPr(plhs[0]);
        ir = mxGetIr(plhs[0]);
        jc = mxGetJc(plhs[0]);
        cnt = (int64_t *)mxMalloc((m2[0] * m2[1] + 1) * sizeof(int64_t));
        for (c = 0; c <= m2[0] * m2[1]; ++c) cnt[c] = 0;
        for (j = 0; j < nnz; ++j) cnt[cols[j] + 1]++;
        for (c = 0; c < m2[0] * m2[1]; ++c) cnt[c + 1] += cnt[c];
        for (c = 0; c <= m2[0] * m2[1]; ++c) jc[c] = (mwIndex)cnt[c];
        for (s = 0; s < N; ++s)                       /* rows ascending within each column */
            for (j = row_ptr[s]; j < row_ptr[s + 1]; ++j) {
                const int64_t at = cnt[cols[j]]++;
                ir[at] = (mwIndex)s;
                pr[at] = vals[j];
            }
        mxFree(cnt);
        mxFree(row_ptr);
        mxFree(cols);
        mxFree(vals);
    }
#elif defined(SID_MEX_PSF)
    {
        /* h = sidPSF('create', H [, CAindex]); sidPSF('destroy', h); sidPSF('precision', h, name);
         * sidPSF('kernel', h, k): the neuron-kernel wrapping of reconstruct_S.m:122-128,184-190 on the
         * handle (forward F(convn(X, k, 'same')), backward convn(B(y), k, 'same')); k = [] removes it */
        char cmd[32], prec[16];
        (void)nlhs;
        if (nrhs < 1 || !mxIsChar(prhs[0]) || mxGetString(prhs[0], cmd, sizeof cmd) != 0)
            arg_error("usage: h = sidPSF('create', H [, CAindex]) | sidPSF('destroy', h) | "
                      "sidPSF('precision', h, name) | sidPSF('kernel', h, k)");
        if (strcmp(cmd, "create") == 0) {
            int64_t dims[5];
            int owned, rc;
            float *H;
            int32_t *CA = NULL;
            sid_psf *p = NULL;
            uint64_t v;
            const mxArray *Ha;
            if (nrhs < 2 || nrhs > 3) arg_error("usage: h = sidPSF('create', H [, CAindex])");
            Ha = psf_array(prhs[1]);
            if (mxGetNumberOfDimensions(Ha) > 5) arg_error("H must be [M, M, Nnum, Nnum, Nz]");
            dims_of(Ha, dims, 5);
            H = as_single(Ha, &owned);
            if (nrhs == 3) CA = caindex_of(prhs[2], dims[4]);
            rc = sid_psf_create(H, dims, CA, &p);
            if (owned) mxFree(H);
            if (CA) mxFree(CA);
            if (rc != SID_OK) fail(rc, "sidPSF create");
            v = (uint64_t)(uintptr_t)p;
            plhs[0] = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
            memcpy(mxGetData(plhs[0]), &v, sizeof v);
        } else if (strcmp(cmd, "destroy") == 0) {
            if (nrhs != 2) arg_error("usage: sidPSF('destroy', h)");
            sid_psf_destroy(handle_arg(prhs[1]));
        } else if (strcmp(cmd, "precision") == 0) {
            int pc = -1, rc;
            if (nrhs != 3 || !mxIsChar(prhs[2]) || mxGetString(prhs[2], prec, sizeof prec) != 0)
                arg_error("usage: sidPSF('precision', h, 'fp32' | 'bf16x2' | 'bf16x3' | 'fp64')");
            if (strcmp(prec, "fp32") == 0) pc = SID_PREC_FP32;
            else if (strcmp(prec, "bf16x2") == 0) pc = SID_PREC_BF16X2;
            else if (strcmp(prec, "bf16x3") == 0) pc = SID_PREC_BF16X3;
            else if (strcmp(prec, "fp64") == 0) pc = SID_PREC_FP64;
            else arg_error("precision must be 'fp32', 'bf16x2', 'bf16x3' or 'fp64'");
            rc = sid_psf_set_precision(handle_arg(prhs[1]), pc);
            if (rc != SID_OK) fail(rc, "sidPSF precision");
        } else if (strcmp(cmd, "kernel") == 0) {
            sid_psf *p;
            int rc, owned = 0;
            if (nrhs != 3 || !is_real_numeric(prhs[2]))
                arg_error("usage: sidPSF('kernel', h, k) with k a real [kx, ky, kz] array, or [] to remove it");
            p = handle_arg(prhs[1]);
            if (mxGetNumberOfElements(prhs[2]) == 0) {
                rc = sid_psf_set_kernel(p, NULL, NULL);
            } else {
                int64_t kd[3];
                float *k;
                if (mxGetNumberOfDimensions(prhs[2]) > 3) arg_error("the kernel must be [kx, ky, kz]");
                dims_of(prhs[2], kd, 3);
                k = as_single(prhs[2], &owned);
                rc = sid_psf_set_kernel(p, k, kd);
                if (owned) mxFree(k);
            }
            if (rc != SID_OK) fail(rc, "sidPSF kernel");
        } else {
            arg_error("sidPSF: unknown command");
        }
    }
#else
#error "define SID_MEX_FORWARD, SID_MEX_BACKWARD, SID_MEX_ISRA, SID_MEX_LS, SID_MEX_LIBGPU, SID_MEX_LIBCPU or SID_MEX_PSF"
#endif
}
