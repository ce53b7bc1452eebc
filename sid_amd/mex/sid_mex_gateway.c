/*
 * MEX gateways over libsidproj (include/sid_proj.h) with the reference's
 * MATLAB names and argument order.  One source, built once per entry point:
 *
 *   mex -DSID_MEX_FORWARD  -output forwardProjectGPU      sid_mex_gateway.c -lsidproj
 *   mex -DSID_MEX_BACKWARD -output backwardProjectGPU_new sid_mex_gateway.c -lsidproj
 *   mex -DSID_MEX_FORWARD  -DSID_MEX_CROP -output forwardProjectACC      sid_mex_gateway.c -lsidproj
 *   mex -DSID_MEX_BACKWARD -DSID_MEX_CROP -output backwardProjectACC_new sid_mex_gateway.c -lsidproj
 *   mex -DSID_MEX_ISRA     -output sidDeconvRL            sid_mex_gateway.c -lsidproj
 *
 * Replaces: external/LFrecon_util/forwardProjectGPU.m:1 (O = forwardProjectGPU(H, realspace)),
 * backwardProjectGPU_new.m:1 (B = backwardProjectGPU_new(H, projection)), forwardProjectACC.m:1 and
 * backwardProjectACC_new.m:1 (third argument CAindex), and the inner loop of deconvRL.m:46-93
 * (sidDeconvRL(H, Htf, maxIter, Xguess) = deconvRL(@(x) forwardProjectGPU(H,x),
 * @(y) backwardProjectGPU_new(H,y), Htf, maxIter, Xguess) with opts.mode 'basic', lambda 0).
 *
 * H, volumes and images arrive as MATLAB column-major arrays: exactly the layout the C ABI takes.
 * Double inputs are converted to single (forwardProjectGPU.m:6,21 computes in single).  The PSF
 * handle (H resident on the GPU with its projector weights) is cached between calls, keyed on the
 * H array's data pointer, dimensions and a strided checksum; it is released at mexAtExit.
 * Requires MATLAB's mex.h; not built in this repository (no MATLAB here).
 */
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "mex.h"
#include "sid_proj.h"

static sid_psf *g_psf = NULL;
static const void *g_key_ptr = NULL;
static int64_t g_key_dims[5];
static double g_key_sum = 0.0;
static int g_key_crop = 0;

static void release(void)
{
    if (g_psf) sid_psf_destroy(g_psf);
    g_psf = NULL;
    g_key_ptr = NULL;
}

static void fail(int rc, const char *where)
{
    char id[64];
    static const char *names[] = {"ok", "arg", "dims", "nz", "oom", "hip", "state", "nodev"};
    snprintf(id, sizeof id, "sid:%s", (rc >= 0 && rc < 8) ? names[rc] : "unknown");
    mexErrMsgIdAndTxt(id, "%s: %s", where, sid_last_error());
}

/* single-precision copy of a numeric array (mxMalloc'ed when converted) */
static float *as_single(const mxArray *a, int *owned)
{
    const mwSize n = mxGetNumberOfElements(a);
    if (mxIsSingle(a)) {
        *owned = 0;
        return (float *)mxGetData(a);
    }
    if (!mxIsDouble(a) || mxIsComplex(a)) mexErrMsgIdAndTxt("sid:arg", "expected a real single or double array");
    {
        const double *d = mxGetPr(a);
        float *f = (float *)mxMalloc(n * sizeof(float));
        mwSize i;
        for (i = 0; i < n; ++i) f[i] = (float)d[i];
        *owned = 1;
        return f;
    }
}

static void dims_of(const mxArray *a, int64_t *out, int k)
{
    const mwSize nd = mxGetNumberOfDimensions(a);
    const mwSize *d = mxGetDimensions(a);
    int i;
    for (i = 0; i < k; ++i) out[i] = (i < (int)nd) ? (int64_t)d[i] : 1;
}

static sid_psf *psf_for(const mxArray *Ha, const mxArray *CAa)
{
    int64_t dims[5];
    double sum = 0.0;
    const int crop = CAa != NULL;
    int owned, rc;
    float *H;
    mwSize n, i, step;
    dims_of(Ha, dims, 5);
    n = mxGetNumberOfElements(Ha);
    step = n / 4096 + 1;
    H = as_single(Ha, &owned);
    for (i = 0; i < n; i += step) sum += (double)H[i];
    if (g_psf && g_key_ptr == mxGetData(Ha) && memcmp(g_key_dims, dims, sizeof dims) == 0 && g_key_sum == sum &&
        g_key_crop == crop) {
        if (owned) mxFree(H);
        return g_psf;
    }
    release();
    {
        int32_t *CA = NULL;
        if (crop) {
            const mwSize m = mxGetNumberOfElements(CAa);
            int ca_owned;
            float *caf = as_single(CAa, &ca_owned);
            CA = (int32_t *)mxMalloc(m * sizeof(int32_t));
            for (i = 0; i < m; ++i) CA[i] = (int32_t)caf[i];
            if (ca_owned) mxFree(caf);
        }
        rc = sid_psf_create(H, dims, CA, &g_psf);
        if (CA) mxFree(CA);
    }
    if (owned) mxFree(H);
    if (rc != SID_OK) fail(rc, "sid_psf_create");
    g_key_ptr = mxGetData(Ha);
    memcpy(g_key_dims, dims, sizeof dims);
    g_key_sum = sum;
    g_key_crop = crop;
    mexAtExit(release);
    return g_psf;
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs;
#if defined(SID_MEX_FORWARD)
    {
        /* O = forwardProjectGPU(H, realspace)  /  forwardProjectACC(H, realspace, CAindex) */
        int64_t vd[3];
        int owned, rc;
        float *V;
        mwSize od[2];
        sid_psf *p;
#ifdef SID_MEX_CROP
        if (nrhs != 3) mexErrMsgIdAndTxt("sid:arg", "usage: O = forwardProjectACC(H, realspace, CAindex)");
        p = psf_for(prhs[0], prhs[2]);
#else
        if (nrhs != 2) mexErrMsgIdAndTxt("sid:arg", "usage: O = forwardProjectGPU(H, realspace)");
        p = psf_for(prhs[0], NULL);
#endif
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
        /* B = backwardProjectGPU_new(H, projection)  /  backwardProjectACC_new(H, projection, CAindex) */
        int64_t odims[2];
        int64_t hd[5];
        int owned, rc;
        float *O;
        mwSize bd[3];
        sid_psf *p;
#ifdef SID_MEX_CROP
        if (nrhs != 3) mexErrMsgIdAndTxt("sid:arg", "usage: B = backwardProjectACC_new(H, projection, CAindex)");
        p = psf_for(prhs[0], prhs[2]);
#else
        if (nrhs != 2) mexErrMsgIdAndTxt("sid:arg", "usage: B = backwardProjectGPU_new(H, projection)");
        p = psf_for(prhs[0], NULL);
#endif
        dims_of(prhs[0], hd, 5);
        dims_of(prhs[1], odims, 2);
        O = as_single(prhs[1], &owned);
        bd[0] = (mwSize)odims[0];
        bd[1] = (mwSize)odims[1];
        bd[2] = (mwSize)hd[4];
        plhs[0] = mxCreateNumericArray(3, bd, mxSINGLE_CLASS, mxREAL);
        rc = sid_backward(p, O, odims, (float *)mxGetData(plhs[0]));
        if (owned) mxFree(O);
        if (rc != SID_OK) fail(rc, "backwardProjectGPU_new");
    }
#elif defined(SID_MEX_ISRA)
    {
        /* X = sidDeconvRL(H, Htf, maxIter, Xguess) */
        int64_t vd[3];
        int oh, ox, rc;
        float *Htf, *X0;
        mwSize xd[3];
        sid_psf *p;
        if (nrhs != 4) mexErrMsgIdAndTxt("sid:arg", "usage: X = sidDeconvRL(H, Htf, maxIter, Xguess)");
        p = psf_for(prhs[0], NULL);
        dims_of(prhs[1], vd, 3);
        Htf = as_single(prhs[1], &oh);
        X0 = as_single(prhs[3], &ox);
        xd[0] = (mwSize)vd[0];
        xd[1] = (mwSize)vd[1];
        xd[2] = (mwSize)vd[2];
        plhs[0] = mxCreateNumericArray(3, xd, mxSINGLE_CLASS, mxREAL);
        memcpy(mxGetData(plhs[0]), X0, (size_t)(vd[0] * vd[1] * vd[2]) * sizeof(float));
        rc = sid_deconv_rl(p, Htf, (float *)mxGetData(plhs[0]), vd, (int)mxGetScalar(prhs[2]), NULL);
        if (oh) mxFree(Htf);
        if (ox) mxFree(X0);
        if (rc != SID_OK) fail(rc, "sidDeconvRL");
    }
#else
#error "define SID_MEX_FORWARD, SID_MEX_BACKWARD or SID_MEX_ISRA"
#endif
}
