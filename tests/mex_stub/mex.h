/*
 * Test stand-in for the part of MATLAB's MEX API that sid_mex_gateway.c uses.
 * TEST INFRASTRUCTURE ONLY: it lets the gateway be compiled and called from the
 * test suite (tests/test_mex.py, tests/test_gpu_mex.py) without MATLAB; the
 * real gateway is built with MATLAB's own mex.h.
 */
#ifndef SID_TEST_MEX_H
#define SID_TEST_MEX_H
#include <stddef.h>

typedef size_t mwSize;
typedef size_t mwIndex;
typedef enum {
    mxUNKNOWN_CLASS = 0, mxCELL_CLASS, mxSTRUCT_CLASS, mxLOGICAL_CLASS, mxCHAR_CLASS, mxVOID_CLASS,
    mxDOUBLE_CLASS, mxSINGLE_CLASS, mxINT8_CLASS, mxUINT8_CLASS, mxINT16_CLASS, mxUINT16_CLASS,
    mxINT32_CLASS, mxUINT32_CLASS, mxINT64_CLASS, mxUINT64_CLASS
} mxClassID;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;
typedef struct mxArray_tag mxArray;

mwSize mxGetNumberOfElements(const mxArray *a);
mwSize mxGetNumberOfDimensions(const mxArray *a);
const mwSize *mxGetDimensions(const mxArray *a);
size_t mxGetM(const mxArray *a);
size_t mxGetN(const mxArray *a);
int mxIsSingle(const mxArray *a);
int mxIsDouble(const mxArray *a);
int mxIsUint64(const mxArray *a);
int mxIsComplex(const mxArray *a);
int mxIsSparse(const mxArray *a);
int mxIsStruct(const mxArray *a);
int mxIsCell(const mxArray *a);
int mxIsChar(const mxArray *a);
void *mxGetData(const mxArray *a);
double *mxGetPr(const mxArray *a);
double mxGetScalar(const mxArray *a);
mxArray *mxGetField(const mxArray *a, mwIndex i, const char *name);
mxArray *mxGetCell(const mxArray *a, mwIndex i);
int mxGetString(const mxArray *a, char *buf, mwSize len);
mwIndex *mxGetIr(const mxArray *a);
mwIndex *mxGetJc(const mxArray *a);
mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID cls, mxComplexity cx);
mxArray *mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity cx);
mxArray *mxCreateSparse(mwSize m, mwSize n, mwSize nzmax, mxComplexity cx);
void *mxMalloc(size_t n);
void mxFree(void *p);
void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...);
mxArray *mxCreateString(const char *s);
int mexCallMATLAB(int nlhs, mxArray *plhs[], int nrhs, mxArray *prhs[], const char *name);
int mexAtExit(void (*fn)(void));

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);
#endif
