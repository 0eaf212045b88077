/* gpu/mxGPUArray.h (test stub) -- declarations of the MATLAB GPU API
 * functions the wrappers call, written for this repository's tests (not
 * MathWorks' header); see ../mex.h.  The stub's "GPU" arrays are host
 * arrays flagged as such, enough to drive the wrappers' gpuArray branches
 * and to count the mxGPUArray views they create and destroy. */
#ifndef FPM_TEST_MXGPUARRAY_STUB_H
#define FPM_TEST_MXGPUARRAY_STUB_H
#include "../mex.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mxGPUArray_tag mxGPUArray;
typedef enum { MX_GPU_INITIALIZE_VALUES = 0, MX_GPU_DO_NOT_INITIALIZE = 1 } mxGPUInitialize;

int mxInitGPU(void);
mxGPUArray const* mxGPUCreateFromMxArray(mxArray const* a);
void const* mxGPUGetDataReadOnly(mxGPUArray const* g);
void* mxGPUGetData(mxGPUArray* g);
mwSize const* mxGPUGetDimensions(mxGPUArray const* g);
mwSize mxGPUGetNumberOfDimensions(mxGPUArray const* g);
mxClassID mxGPUGetClassID(mxGPUArray const* g);
mxComplexity mxGPUGetComplexity(mxGPUArray const* g);
mxGPUArray* mxGPUCreateGPUArray(mwSize nd, mwSize const* dims, mxClassID cls, mxComplexity cx,
                                mxGPUInitialize init);
mxArray* mxGPUCreateMxArrayOnGPU(mxGPUArray const* g);
void mxGPUDestroyGPUArray(mxGPUArray const* g);

#ifdef __cplusplus
}
#endif
#endif
