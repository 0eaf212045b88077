// mex_stub.cpp -- host implementation of the MATLAB API declared in mex.h and
// gpu/mxGPUArray.h (test stubs), plus a small C harness (stub_*) through
// which the tests build MATLAB-like arrays, call a wrapper's mexFunction and
// read back its outputs, the error it raised (mexErrMsgIdAndTxt throws here)
// and how many arrays and mxGPUArray views are still alive.
#include "mex.h"
#include "gpu/mxGPUArray.h"

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

struct mxArray_tag {
    std::vector<mwSize> dims;
    mxClassID cls = mxUNKNOWN_CLASS;
    bool complex = false;
    bool gpu = false;  // flagged as a gpuArray (data still in host memory)
    std::vector<char> data;
};

struct mxGPUArray_tag {
    mxArray* arr = nullptr;  // owned
};

namespace {

int g_live_arrays = 0, g_live_views = 0, g_init_gpu = 0;
std::string g_err_id, g_err_msg;
struct MexError {};

size_t elem_bytes(mxClassID c, bool cx) {
    size_t b = c == mxDOUBLE_CLASS ? 8 : c == mxSINGLE_CLASS ? 4 : c == mxINT32_CLASS ? 4 : c == mxUINT16_CLASS ? 2 : 1;
    return cx ? 2 * b : b;
}

size_t numel(const mxArray* a) {
    size_t n = 1;
    for (mwSize d : a->dims) n *= d;
    return n;
}

mxArray* make(size_t nd, const mwSize* dims, mxClassID cls, bool cx, bool gpu) {
    mxArray* a = new mxArray_tag;
    a->dims.assign(dims, dims + nd);
    a->cls = cls;
    a->complex = cx;
    a->gpu = gpu;
    a->data.assign(numel(a) * elem_bytes(cls, cx), 0);
    ++g_live_arrays;
    return a;
}

}  // namespace

extern "C" {

int mxIsGPUArray(const mxArray* a) { return a->gpu ? 1 : 0; }
const mwSize* mxGetDimensions(const mxArray* a) { return a->dims.data(); }
mwSize mxGetNumberOfDimensions(const mxArray* a) { return a->dims.size(); }
size_t mxGetNumberOfElements(const mxArray* a) { return numel(a); }
mxClassID mxGetClassID(const mxArray* a) { return a->cls; }
int mxIsComplex(const mxArray* a) { return a->complex ? 1 : 0; }
int mxIsDouble(const mxArray* a) { return a->cls == mxDOUBLE_CLASS ? 1 : 0; }
void* mxGetData(const mxArray* a) { return (void*)a->data.data(); }
mxComplexSingle* mxGetComplexSingles(const mxArray* a) {
    return a->cls == mxSINGLE_CLASS && a->complex ? (mxComplexSingle*)a->data.data() : nullptr;
}
double* mxGetDoubles(const mxArray* a) {
    return a->cls == mxDOUBLE_CLASS && !a->complex ? (double*)a->data.data() : nullptr;
}
double mxGetScalar(const mxArray* a) {
    if (a->data.empty()) return 0.0;
    switch (a->cls) {
        case mxDOUBLE_CLASS: return *(const double*)a->data.data();
        case mxSINGLE_CLASS: return *(const float*)a->data.data();
        case mxINT32_CLASS: return *(const int*)a->data.data();
        case mxUINT16_CLASS: return *(const unsigned short*)a->data.data();
        default: return 0.0;
    }
}
mxArray* mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity cx) {
    const mwSize d[2] = {m, n};
    return make(2, d, cls, cx == mxCOMPLEX, false);
}
void mxDestroyArray(mxArray* a) {
    if (a) {
        --g_live_arrays;
        delete a;
    }
}
void mxFree(void* p) { std::free(p); }
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err_id = id;
    g_err_msg = buf;
    throw MexError{};
}

int mxInitGPU(void) {
    ++g_init_gpu;
    return 0;
}
mxGPUArray const* mxGPUCreateFromMxArray(mxArray const* a) {
    mxGPUArray* g = new mxGPUArray_tag;
    g->arr = make(a->dims.size(), a->dims.data(), a->cls, a->complex, true);
    g->arr->data = a->data;
    ++g_live_views;
    return g;
}
void const* mxGPUGetDataReadOnly(mxGPUArray const* g) { return g->arr->data.data(); }
void* mxGPUGetData(mxGPUArray* g) { return g->arr->data.data(); }
mwSize const* mxGPUGetDimensions(mxGPUArray const* g) {
    mwSize* d = (mwSize*)std::malloc(sizeof(mwSize) * g->arr->dims.size());
    std::memcpy(d, g->arr->dims.data(), sizeof(mwSize) * g->arr->dims.size());
    return d;
}
mwSize mxGPUGetNumberOfDimensions(mxGPUArray const* g) { return g->arr->dims.size(); }
mxClassID mxGPUGetClassID(mxGPUArray const* g) { return g->arr->cls; }
mxComplexity mxGPUGetComplexity(mxGPUArray const* g) { return g->arr->complex ? mxCOMPLEX : mxREAL; }
mxGPUArray* mxGPUCreateGPUArray(mwSize nd, mwSize const* dims, mxClassID cls, mxComplexity cx, mxGPUInitialize) {
    mxGPUArray* g = new mxGPUArray_tag;
    g->arr = make(nd, dims, cls, cx == mxCOMPLEX, true);
    ++g_live_views;
    return g;
}
mxArray* mxGPUCreateMxArrayOnGPU(mxGPUArray const* g) {
    mxArray* a = make(g->arr->dims.size(), g->arr->dims.data(), g->arr->cls, g->arr->complex, true);
    a->data = g->arr->data;
    return a;
}
void mxGPUDestroyGPUArray(mxGPUArray const* g) {
    if (!g) return;
    --g_live_views;
    mxDestroyArray(g->arr);
    delete g;
}

// ------------------------------------------------------------- harness
mxArray* stub_create(int cls, int complex, int nd, const size_t* dims, const void* data, int gpu) {
    std::vector<mwSize> d(dims, dims + nd);
    mxArray* a = make((size_t)nd, d.data(), (mxClassID)cls, complex != 0, gpu != 0);
    if (data) std::memcpy(a->data.data(), data, a->data.size());
    return a;
}
void* stub_data(mxArray* a) { return a->data.data(); }
size_t stub_bytes(mxArray* a) { return a->data.size(); }
int stub_ndims(mxArray* a) { return (int)a->dims.size(); }
size_t stub_dim(mxArray* a, int k) { return a->dims[k]; }
int stub_class(mxArray* a) { return (int)a->cls; }
void stub_destroy(mxArray* a) { mxDestroyArray(a); }
int stub_call(int nlhs, mxArray** plhs, int nrhs, mxArray** prhs) {
    g_err_id.clear();
    g_err_msg.clear();
    try {
        mexFunction(nlhs, plhs, nrhs, (const mxArray**)prhs);
    } catch (const MexError&) {
        return 1;
    }
    return 0;
}
const char* stub_error_id(void) { return g_err_id.c_str(); }
const char* stub_error_msg(void) { return g_err_msg.c_str(); }
int stub_live_arrays(void) { return g_live_arrays; }
int stub_live_views(void) { return g_live_views; }
int stub_init_gpu_calls(void) { return g_init_gpu; }
int stub_class_id(const char* name) {
    const std::string n = name;
    return n == "double" ? mxDOUBLE_CLASS : n == "single" ? mxSINGLE_CLASS : n == "int32" ? mxINT32_CLASS
         : n == "uint16" ? mxUINT16_CLASS : mxUNKNOWN_CLASS;
}

}  // extern "C"
