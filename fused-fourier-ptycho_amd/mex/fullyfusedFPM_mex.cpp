// fullyfusedFPM_mex.cpp -- optional MATLAB wrapper: the gradient-only mex of
// benchmark/cuFPM/fullyfusedFPM.cu:21-162 on top of libfpm_amd
// (fpm_fused_gradient, with the loss of cuFPM_v3/cuda/kernels.cuh:197-275).
//
//   [dldw1, dldw2] = fullyfusedFPM(w1, w2, Y, ledIdx, pratio)
//
// so that the reference's own MATLAB loop (codes/main_run_FPM.m:106-160:
// gradient, optimizer_w1.step, optimizer_w2.step, pupil constraint) runs
// unchanged on these kernels.  Y holds the minibatch's pages (single
// amplitudes prepared once over the whole stack, as that loop does); a
// uint16 stack is also accepted, but is then normalised over the pages
// passed.  Built by the user against their own MATLAB installation (mex.h and
// the GPU API header are MathWorks'; they are not part of this repository):
//   mex -R2018a fullyfusedFPM_mex.cpp -I<repo>/include -L<lib dir> -lfpm_amd -output fullyfusedFPM
// gpuArray inputs are passed as device pointers (outputs are gpuArrays);
// host arrays are staged by the library.  Errors are raised as MATLAB errors
// with the library's text.
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

}  // namespace

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
    if (nrhs != 5) mexErrMsgIdAndTxt("fpm:nrhs", "fullyfusedFPM takes 5 inputs (w1, w2, Y, ledIdx, pratio)");
    if (nlhs > 2) mexErrMsgIdAndTxt("fpm:nlhs", "fullyfusedFPM returns at most 2 outputs");
    // MATLAB's GPU library (gpuArray) exists only on CUDA devices: on an
    // MI355X host every input is a host array and mxInitGPU is not called
    bool any_gpu = false;
    for (int k = 0; k < 4; ++k) any_gpu = any_gpu || mxIsGPUArray(prhs[k]);
    if (any_gpu) mxInitGPU();
    In w1 = view(prhs[0]), w2 = view(prhs[1]), Y = view(prhs[2]), led = view(prhs[3]);
    const In* all[] = {&w1, &w2, &Y, &led};
    if (w1.cls != mxSINGLE_CLASS || !w1.complex || w2.cls != mxSINGLE_CLASS || !w2.complex) {
        release(all, 4);
        mexErrMsgIdAndTxt("fpm:type", "w1 and w2 must be complex single");
    }
    if (led.cls != mxINT32_CLASS) {
        release(all, 4);
        mexErrMsgIdAndTxt("fpm:type", "ledIdx must be int32 (2 x L)");
    }
    const int32_t ytype = Y.cls == mxUINT16_CLASS ? FPM_Y_U16_RAW : FPM_Y_F32_AMPLITUDE;
    const bool dev = w1.g && w2.g && Y.g && led.g;
    if (!dev && (w1.g || w2.g || Y.g || led.g)) {
        release(all, 4);
        mexErrMsgIdAndTxt("fpm:mixed", "pass either all arrays as gpuArray or none");
    }
    fpm_opts o;
    fpm_opts_init(&o);
    o.pointers_are_device = dev ? 1 : 0;

    mxArray *G1 = nullptr, *G2 = nullptr;
    mxGPUArray *G1g = nullptr, *G2g = nullptr;
    void *g1p, *g2p;
    if (dev) {
        const mwSize d1[2] = {w1.rows, w1.cols}, d2[2] = {w2.rows, w2.cols};
        G1g = mxGPUCreateGPUArray(2, d1, mxSINGLE_CLASS, mxCOMPLEX, MX_GPU_DO_NOT_INITIALIZE);
        G2g = mxGPUCreateGPUArray(2, d2, mxSINGLE_CLASS, mxCOMPLEX, MX_GPU_DO_NOT_INITIALIZE);
        g1p = mxGPUGetData(G1g);
        g2p = mxGPUGetData(G2g);
    } else {
        G1 = mxCreateNumericMatrix(w1.rows, w1.cols, mxSINGLE_CLASS, mxCOMPLEX);
        G2 = mxCreateNumericMatrix(w2.rows, w2.cols, mxSINGLE_CLASS, mxCOMPLEX);
        g1p = mxGetComplexSingles(G1);
        g2p = mxGetComplexSingles(G2);
    }
    const int rc = fpm_fused_gradient((const fpm_c32*)w1.ptr, (int64_t)w1.rows, (int64_t)w1.cols,
                                      (const fpm_c32*)w2.ptr, (int64_t)w2.rows, (int64_t)w2.cols, Y.ptr, ytype,
                                      (int64_t)Y.pages, (const int32_t*)led.ptr, mxGetScalar(prhs[4]),
                                      (fpm_c32*)g1p, (fpm_c32*)g2p, nullptr, &o);
    release(all, 4);
    if (rc != FPM_OK) {
        if (G1g) mxGPUDestroyGPUArray(G1g);
        if (G2g) mxGPUDestroyGPUArray(G2g);
        if (G1) mxDestroyArray(G1);
        if (G2) mxDestroyArray(G2);
        mexErrMsgIdAndTxt("fpm:gradient", "%s", fpm_last_error());
    }
    if (dev) {
        plhs[0] = mxGPUCreateMxArrayOnGPU(G1g);
        plhs[1] = mxGPUCreateMxArrayOnGPU(G2g);
        mxGPUDestroyGPUArray(G1g);
        mxGPUDestroyGPUArray(G2g);
    } else {
        plhs[0] = G1;
        plhs[1] = G2;
    }
}
