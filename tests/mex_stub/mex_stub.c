/*
 * Test stand-in for MATLAB's mxArray / MEX runtime (TEST INFRASTRUCTURE ONLY),
 * linked into each build of sid_mex_gateway.c so the tests can call its
 * mexFunction through ctypes.  Arrays are column-major like MATLAB's;
 * mexErrMsgIdAndTxt records the error identifier and message and returns
 * control to stub_call through longjmp (MATLAB unwinds to its prompt).
 */
#include <setjmp.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

struct mxArray_tag {
    mxClassID cls;
    int sparse;
    mwSize ndim;
    mwSize dims[8];
    void *data;
    mwIndex *ir, *jc;
    mwSize nzmax;
    int nfields;
    char **fnames;
    mxArray **fvals;
    mxArray **cells;
};

static size_t elem_size(mxClassID c)
{
    switch (c) {
        case mxDOUBLE_CLASS: case mxINT64_CLASS: case mxUINT64_CLASS: return 8;
        case mxSINGLE_CLASS: case mxINT32_CLASS: case mxUINT32_CLASS: return 4;
        case mxINT16_CLASS: case mxUINT16_CLASS: case mxCHAR_CLASS: return 2;
        default: return 1;
    }
}

mwSize mxGetNumberOfElements(const mxArray *a)
{
    mwSize n = 1, i;
    for (i = 0; i < a->ndim; ++i) n *= a->dims[i];
    return n;
}
mwSize mxGetNumberOfDimensions(const mxArray *a) { return a->ndim; }
const mwSize *mxGetDimensions(const mxArray *a) { return a->dims; }
size_t mxGetM(const mxArray *a) { return a->dims[0]; }
size_t mxGetN(const mxArray *a)
{
    size_t n = 1, i;
    for (i = 1; i < a->ndim; ++i) n *= a->dims[i];
    return n;
}
int mxIsSingle(const mxArray *a) { return a->cls == mxSINGLE_CLASS; }
int mxIsDouble(const mxArray *a) { return a->cls == mxDOUBLE_CLASS; }
int mxIsUint64(const mxArray *a) { return a->cls == mxUINT64_CLASS; }
int mxIsComplex(const mxArray *a) { (void)a; return 0; }
int mxIsSparse(const mxArray *a) { return a->sparse; }
int mxIsStruct(const mxArray *a) { return a->cls == mxSTRUCT_CLASS; }
int mxIsCell(const mxArray *a) { return a->cls == mxCELL_CLASS; }
int mxIsChar(const mxArray *a) { return a->cls == mxCHAR_CLASS; }
void *mxGetData(const mxArray *a) { return a->data; }
double *mxGetPr(const mxArray *a) { return (double *)a->data; }
double mxGetScalar(const mxArray *a)
{
    switch (a->cls) {
        case mxDOUBLE_CLASS: return ((double *)a->data)[0];
        case mxSINGLE_CLASS: return ((float *)a->data)[0];
        case mxINT32_CLASS: return ((int32_t *)a->data)[0];
        case mxUINT64_CLASS: return (double)((uint64_t *)a->data)[0];
        default: return 0.0;
    }
}
mxArray *mxGetField(const mxArray *a, mwIndex i, const char *name)
{
    int f;
    if (a->cls != mxSTRUCT_CLASS || i != 0) return NULL;
    for (f = 0; f < a->nfields; ++f)
        if (strcmp(a->fnames[f], name) == 0) return a->fvals[f];
    return NULL;
}
mxArray *mxGetCell(const mxArray *a, mwIndex i) { return a->cls == mxCELL_CLASS ? a->cells[i] : NULL; }
int mxGetString(const mxArray *a, char *buf, mwSize len)
{
    mwSize n = mxGetNumberOfElements(a), i;
    if (a->cls != mxCHAR_CLASS || n + 1 > len) return 1;
    for (i = 0; i < n; ++i) buf[i] = (char)((uint16_t *)a->data)[i];
    buf[n] = 0;
    return 0;
}
mwIndex *mxGetIr(const mxArray *a) { return a->ir; }
mwIndex *mxGetJc(const mxArray *a) { return a->jc; }

static mxArray *new_array(mxClassID cls, mwSize ndim, const mwSize *dims)
{
    mxArray *a = (mxArray *)calloc(1, sizeof(mxArray));
    mwSize i, n;
    a->cls = cls;
    a->ndim = ndim < 2 ? 2 : ndim;
    a->dims[0] = a->dims[1] = 1;
    for (i = 0; i < ndim && i < 8; ++i) a->dims[i] = dims[i];
    n = mxGetNumberOfElements(a);
    a->data = calloc(n ? n : 1, elem_size(cls));
    return a;
}
mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID cls, mxComplexity cx)
{
    (void)cx;
    return new_array(cls, ndim, dims);
}
mxArray *mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity cx)
{
    mwSize d[2] = {m, n};
    (void)cx;
    return new_array(cls, 2, d);
}
mxArray *mxCreateSparse(mwSize m, mwSize n, mwSize nzmax, mxComplexity cx)
{
    mxArray *a = (mxArray *)calloc(1, sizeof(mxArray));
    (void)cx;
    a->cls = mxDOUBLE_CLASS;
    a->sparse = 1;
    a->ndim = 2;
    a->dims[0] = m;
    a->dims[1] = n;
    a->nzmax = nzmax ? nzmax : 1;
    a->data = calloc(a->nzmax, sizeof(double));
    a->ir = (mwIndex *)calloc(a->nzmax, sizeof(mwIndex));
    a->jc = (mwIndex *)calloc(n + 1, sizeof(mwIndex));
    return a;
}
void *mxMalloc(size_t n) { return malloc(n ? n : 1); }
void mxFree(void *p) { free(p); }

static jmp_buf g_jmp;
static char g_err_id[128], g_err_msg[1024];
void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...)
{
    va_list ap;
    snprintf(g_err_id, sizeof g_err_id, "%s", id);
    va_start(ap, fmt);
    vsnprintf(g_err_msg, sizeof g_err_msg, fmt, ap);
    va_end(ap);
    longjmp(g_jmp, 1);
}
static void (*g_exit_fn)(void) = NULL;
int mexAtExit(void (*fn)(void))
{
    g_exit_fn = fn;
    return 0;
}

/* mexCallMATLAB: a MATLAB function the gateway calls back (generate_kernel);
 * the harness registers one handler, which returns 0 and fills plhs, or
 * nonzero when the function is unknown (MATLAB then raises its own error). */
typedef int (*stub_matlab_fn)(int nlhs, mxArray **plhs, int nrhs, mxArray **prhs, const char *name);
static stub_matlab_fn g_matlab = NULL;
void stub_set_matlab(stub_matlab_fn fn) { g_matlab = fn; }
int mexCallMATLAB(int nlhs, mxArray *plhs[], int nrhs, mxArray *prhs[], const char *name)
{
    if (!g_matlab || g_matlab(nlhs, plhs, nrhs, prhs, name) != 0)
        mexErrMsgIdAndTxt("MATLAB:UndefinedFunction", "Undefined function '%s'.", name);
    return 0;
}

/* ---- harness entry points (called from Python through ctypes) ---- */
mxArray *stub_numeric(int cls, int ndim, const int64_t *dims, const void *data)
{
    mwSize d[8];
    int i;
    mxArray *a;
    for (i = 0; i < ndim && i < 8; ++i) d[i] = (mwSize)dims[i];
    a = new_array((mxClassID)cls, (mwSize)ndim, d);
    if (data) memcpy(a->data, data, mxGetNumberOfElements(a) * elem_size(a->cls));
    return a;
}
mxArray *stub_char(const char *s)
{
    mwSize d[2] = {1, strlen(s)}, i;
    mxArray *a = new_array(mxCHAR_CLASS, 2, d);
    for (i = 0; i < d[1]; ++i) ((uint16_t *)a->data)[i] = (uint16_t)(unsigned char)s[i];
    return a;
}
mxArray *mxCreateString(const char *s) { return stub_char(s); }
mxArray *stub_struct(int n, const char **names, mxArray **vals)
{
    mwSize d[2] = {1, 1};
    mxArray *a = new_array(mxSTRUCT_CLASS, 2, d);
    int f;
    a->nfields = n;
    a->fnames = (char **)calloc(n ? n : 1, sizeof(char *));
    a->fvals = (mxArray **)calloc(n ? n : 1, sizeof(mxArray *));
    for (f = 0; f < n; ++f) {
        a->fnames[f] = (char *)malloc(strlen(names[f]) + 1);
        strcpy(a->fnames[f], names[f]);
        a->fvals[f] = vals[f];
    }
    return a;
}
mxArray *stub_cell(int n, mxArray **elems)
{
    mwSize d[2] = {1, (mwSize)n};
    mxArray *a = new_array(mxCELL_CLASS, 2, d);
    int i;
    a->cells = (mxArray **)calloc(n ? n : 1, sizeof(mxArray *));
    for (i = 0; i < n; ++i) a->cells[i] = elems[i];
    return a;
}
int stub_class(const mxArray *a) { return (int)a->cls; }
int stub_is_sparse(const mxArray *a) { return a->sparse; }
int stub_ndim(const mxArray *a) { return (int)a->ndim; }
int64_t stub_dim(const mxArray *a, int i) { return (int64_t)a->dims[i]; }
void *stub_data(const mxArray *a) { return a->data; }
mwIndex *stub_ir(const mxArray *a) { return a->ir; }
mwIndex *stub_jc(const mxArray *a) { return a->jc; }
const char *stub_err_id(void) { return g_err_id; }
const char *stub_err_msg(void) { return g_err_msg; }
/* 0: returned normally; 1: mexErrMsgIdAndTxt was called (id / message recorded) */
int stub_call(int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs)
{
    g_err_id[0] = g_err_msg[0] = 0;
    if (setjmp(g_jmp)) return 1;
    mexFunction(nlhs, plhs, nrhs, prhs);
    return 0;
}
void stub_exit(void)
{
    if (g_exit_fn) g_exit_fn();
    g_exit_fn = NULL;
}
