/* mex.h (test stub) -- declarations of the few MATLAB C Matrix API / MEX
 * functions the wrappers in fused-fourier-ptycho_amd/mex/ call, written for
 * this repository's tests (not MathWorks' header).  Names and argument lists
 * follow the public MATLAB External API (R2018a interleaved complex); the
 * enum values and the mxArray layout are this stub's own.  mex_stub.cpp
 * implements them on host memory so that the wrappers can be compiled and
 * run without MATLAB. */
#ifndef FPM_TEST_MEX_STUB_H
#define FPM_TEST_MEX_STUB_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef size_t mwSize;
typedef struct mxArray_tag mxArray;
typedef enum {
    mxUNKNOWN_CLASS = 0,
    mxDOUBLE_CLASS,
    mxSINGLE_CLASS,
    mxINT32_CLASS,
    mxUINT16_CLASS
} mxClassID;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;
typedef struct {
    float real, imag;
} mxComplexSingle;

int mxIsGPUArray(const mxArray* a);
const mwSize* mxGetDimensions(const mxArray* a);
mwSize mxGetNumberOfDimensions(const mxArray* a);
size_t mxGetNumberOfElements(const mxArray* a);
mxClassID mxGetClassID(const mxArray* a);
int mxIsComplex(const mxArray* a);
int mxIsDouble(const mxArray* a);
void* mxGetData(const mxArray* a);
mxComplexSingle* mxGetComplexSingles(const mxArray* a);
double* mxGetDoubles(const mxArray* a);
double mxGetScalar(const mxArray* a);
mxArray* mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity cx);
void mxDestroyArray(mxArray* a);
void mxFree(void* p);
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...);
/* the entry every mex file defines (C linkage, as MATLAB looks it up) */
void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);

#ifdef __cplusplus
}
#endif
#endif
