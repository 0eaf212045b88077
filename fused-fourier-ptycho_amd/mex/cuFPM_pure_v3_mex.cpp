// cuFPM_pure_v3_mex.cpp -- optional MATLAB wrapper: the mex call of
// cuFPM_v3/cuFPM_pure_v3.cu:44-228 on top of libfpm_amd.
//
//   [O, P] = cuFPM_pure_v3(O0, P0, Y, ledIdx, A, pratio, numEpochs, batch, [lr beta eps])
//
// Built by the user against their own MATLAB installation (mex.h and the GPU
// API header are MathWorks'; they are not part of this repository):
//   mex -R2018a cuFPM_pure_v3_mex.cpp -I<repo>/include -L<lib dir> -lfpm_amd -output cuFPM_pure_v3
// gpuArray inputs are passed as device pointers; host arrays are staged by
// the library.  Y may be single amplitudes (as the reference takes) or uint16
// raw counts.  Errors are raised as MATLAB errors with the library's text.
#include "mex.h"
#include "gpu/mxGPUArray.h"

#include "fpm_amd.h"

namespace {

struct In {
    const void* ptr = nullptr;
    const mxGPUArray* g = nullptr;  // non-null for gpuArray inputs
    mwSize rows = 0, cols = 0, pages = 1;
    mxClassID cls = mxUNKNOWN_CLASS;
    bool complex = false;
};

In view(const mxArray* a) {
    In v;
    if (mxIsGPUArray(a)) {
        v.g = mxGPUCreateFromMxArray(a);
        v.ptr = mxGPUGetDataReadOnly(v.g);
        mwSize* d = (mwSize*)mxGPUGetDimensions(v.g);  // a copy: freed with mxFree
        const mwSize nd = mxGPUGetNumberOfDimensions(v.g);
        v.rows = d[0];
        v.cols = nd > 1 ? d[1] : 1;
        v.pages = nd > 2 ? d[2] : 1;
        mxFree(d);
        v.cls = mxGPUGetClassID(v.g);
        v.complex = mxGPUGetComplexity(v.g) == mxCOMPLEX;
    } else {
        const mwSize* d = mxGetDimensions(a);
        const mwSize nd = mxGetNumberOfDimensions(a);
        v.rows = d[0];
        v.cols = nd > 1 ? d[1] : 1;
        v.pages = nd > 2 ? d[2] : 1;
        v.cls = mxGetClassID(a);
        v.complex = mxIsComplex(a);
        v.ptr = v.complex ? (const void*)mxGetComplexSingles(a) : mxGetData(a);
    }
    return v;
}

void release(const In* const* vs, int count) {
    for (int k = 0; k < count; ++k)
        if (vs[k]->g) mxGPUDestroyGPUArray(vs[k]->g);
}

// releases the gpuArray views, then raises the MATLAB error (which does not return)
[[noreturn]] void fail(const In* const* vs, int count, const char* id, const char* msg) {
    release(vs, count);
    mexErrMsgIdAndTxt(id, "%s", msg);
    for (;;) {
    }
}

}  // namespace

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    if (nrhs != 9) mexErrMsgIdAndTxt("fpm:nrhs", "cuFPM_pure_v3 takes 9 inputs");
    if (nlhs > 2) mexErrMsgIdAndTxt("fpm:nlhs", "cuFPM_pure_v3 returns at most 2 outputs");
    // MATLAB's GPU library (gpuArray) exists only on CUDA devices: on an
    // MI355X host every input is a host array and mxInitGPU is not called
    bool any_gpu = false;
    for (int k = 0; k < 5; ++k) any_gpu = any_gpu || mxIsGPUArray(prhs[k]);
    if (any_gpu) mxInitGPU();
    In O0 = view(prhs[0]), P0 = view(prhs[1]), Y = view(prhs[2]), led = view(prhs[3]), A = view(prhs[4]);
    const In* all[] = {&O0, &P0, &Y, &led, &A};
    if (O0.cls != mxSINGLE_CLASS || !O0.complex || P0.cls != mxSINGLE_CLASS || !P0.complex)
        fail(all, 5, "fpm:type", "O0 and P0 must be complex single");
    if (led.cls != mxINT32_CLASS || A.cls != mxSINGLE_CLASS) fail(all, 5, "fpm:type", "ledIdx int32, A single");
    const int32_t ytype = Y.cls == mxUINT16_CLASS ? FPM_Y_U16_RAW : FPM_Y_F32_AMPLITUDE;
    const bool dev = O0.g && P0.g && Y.g && led.g && A.g;
    if (!dev && (O0.g || P0.g || Y.g || led.g || A.g)) fail(all, 5, "fpm:mixed", "pass either all arrays as gpuArray or none");
    if (mxGetNumberOfElements(prhs[8]) < 3 || !mxIsDouble(prhs[8]))
        fail(all, 5, "fpm:opt", "optimizer argument is double [lr beta eps]");
    const double* opt = mxGetDoubles(prhs[8]);

    fpm_opts o;
    fpm_opts_init(&o);
    o.pointers_are_device = dev ? 1 : 0;
    o.verbose = 1;  // the reference prints its parameters (cuFPM_pure_v3.cu:61,66)

    mxArray *Oout = nullptr, *Pout = nullptr;
    void *Op, *Pp;
    mxGPUArray *Og = nullptr, *Pg = nullptr;
    if (dev) {
        const mwSize od[2] = {O0.rows, O0.cols}, pd[2] = {P0.rows, P0.cols};
        Og = mxGPUCreateGPUArray(2, od, mxSINGLE_CLASS, mxCOMPLEX, MX_GPU_DO_NOT_INITIALIZE);
        Pg = mxGPUCreateGPUArray(2, pd, mxSINGLE_CLASS, mxCOMPLEX, MX_GPU_DO_NOT_INITIALIZE);
        Op = mxGPUGetData(Og);
        Pp = mxGPUGetData(Pg);
    } else {
        Oout = mxCreateNumericMatrix(O0.rows, O0.cols, mxSINGLE_CLASS, mxCOMPLEX);
        Pout = mxCreateNumericMatrix(P0.rows, P0.cols, mxSINGLE_CLASS, mxCOMPLEX);
        Op = mxGetComplexSingles(Oout);
        Pp = mxGetComplexSingles(Pout);
    }
    const int rc = fpm_pure_v3((const fpm_c32*)O0.ptr, (int64_t)O0.rows, (int64_t)O0.cols, (const fpm_c32*)P0.ptr,
                               (int64_t)P0.rows, (int64_t)P0.cols, Y.ptr, ytype, (int64_t)Y.pages,
                               (const int32_t*)led.ptr, (const float*)A.ptr, mxGetScalar(prhs[5]),
                               mxGetScalar(prhs[6]), mxGetScalar(prhs[7]), opt, (fpm_c32*)Op, (fpm_c32*)Pp, &o);
    release(all, 5);
    if (rc != FPM_OK) {
        if (Og) mxGPUDestroyGPUArray(Og);
        if (Pg) mxGPUDestroyGPUArray(Pg);
        if (Oout) mxDestroyArray(Oout);
        if (Pout) mxDestroyArray(Pout);
        mexErrMsgIdAndTxt("fpm:solve", "%s", fpm_last_error());
    }
    if (dev) {
        plhs[0] = mxGPUCreateMxArrayOnGPU(Og);
        plhs[1] = mxGPUCreateMxArrayOnGPU(Pg);
        mxGPUDestroyGPUArray(Og);
        mxGPUDestroyGPUArray(Pg);
    } else {
        plhs[0] = Oout;
        plhs[1] = Pout;
    }
}
