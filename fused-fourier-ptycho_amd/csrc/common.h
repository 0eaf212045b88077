// common.h -- shared host/device definitions of libfpm_amd.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <cmath>
#include <string>

namespace fpm {

// FFT engines (see fft_core.h): FP32 butterflies on the vector ALUs, or the
// radix-16 DFTs as split-FP16 matrix-core products with FP32 accumulation
enum { ENG_F32 = 0, ENG_SPLIT16 = 1 };

// Column permutation of a frequency-domain array whose column count C is
// split C = M1 * M2 for the two-pass (four-step) column transform: natural
// column k = k1 + M1 k2 is stored at column k1 M2 + k2.  M2 == 1 is the
// identity (columns short enough for one pass).
struct ColPerm {
    uint32_t M1, M2, log2M1, C;
    __host__ __device__ __forceinline__ uint32_t operator()(uint32_t k) const {
        return (k & (M1 - 1)) * M2 + (k >> log2M1);
    }
};

// longest column transform done in a single strided pass (LDS-bound: a
// 2048-point pass uses 8-row tiles, 147 KiB)
#ifndef FPM_MAX_STRIDED_LEN
#define FPM_MAX_STRIDED_LEN 512
#endif
constexpr int kMaxStridedLen = FPM_MAX_STRIDED_LEN;
// rows per strided tile (16 complex floats = 128 bytes)
constexpr int kStridedRows = 16;
// rows per entry of the window row masks (byte per 16 rows)
constexpr int kMaskRows = 16;

inline int ilog2_host(uint64_t v) {
    int l = 0;
    while ((1ull << l) < v) ++l;
    return l;
}
inline bool is_pow2(int64_t v) { return v > 0 && (v & (v - 1)) == 0; }
// 1/x when x is a power of two (then x * (1/x) scaling equals division by x
// exactly), else 0
inline float exact_inverse_or_zero(float x) {
    int e = 0;
    return (x > 0.0f && std::frexp(x, &e) == 0.5f) ? 1.0f / x : 0.0f;
}

inline ColPerm make_colperm(uint32_t C) {
    ColPerm p;
    p.C = C;
    if (C <= (uint32_t)kMaxStridedLen) {
        p.M1 = C; p.M2 = 1;
    } else {
        int lg = ilog2_host(C);
        p.M1 = 1u << (lg / 2);
        p.M2 = C / p.M1;
    }
    p.log2M1 = ilog2_host(p.M1);
    return p;
}

// Layout of a stack of rows x cols pages: element (r, c) of a page at
//   (r >> sh) * tstride + perm(c) * ld + (r & rmask)
//   column-major: sh = 31, tstride = 0, ld = rows, rmask = ~0   (r < 2^31)
//   8-row tiles:  sh = 3, tstride = 8 cols, ld = 8, rmask = 7, perm = identity
// The tiles keep the 8 rows x cols block of a one-pass column transform
// contiguous in HBM (a 2048-point column pass of 8 rows streams 128 KiB),
// while a pass down one column still reads 64-byte runs.
struct PageLayout {
    uint32_t sh, rmask, ld;
    uint64_t tstride;
    ColPerm perm;
    __host__ __device__ __forceinline__ size_t row_off(uint32_t r) const {
        return (size_t)(r >> sh) * tstride + (r & rmask);
    }
    __host__ __device__ __forceinline__ size_t off(uint32_t r, uint32_t c) const {
        return row_off(r) + (size_t)perm(c) * ld;
    }
    __host__ __device__ __forceinline__ bool tiled() const { return sh != 31; }
};
constexpr int kTileRows = 8;
// per-LED minibatch pages in 8-row tiles where possible (0: always column-major)
#ifndef FPM_TILED_PAGES
#define FPM_TILED_PAGES 1
#endif
constexpr bool kTiledPages = FPM_TILED_PAGES != 0;

inline PageLayout colmajor_layout(uint32_t rows, const ColPerm& p) {
    PageLayout l;
    l.sh = 31;
    l.rmask = 0xffffffffu;
    l.ld = rows;
    l.tstride = 0;
    l.perm = p;
    return l;
}

inline PageLayout tiled_layout(uint32_t cols) {
    PageLayout l;
    l.sh = 3;
    l.rmask = kTileRows - 1;
    l.ld = kTileRows;
    l.tstride = (uint64_t)kTileRows * cols;
    l.perm.C = cols;
    l.perm.M1 = cols;
    l.perm.M2 = 1;
    l.perm.log2M1 = ilog2_host(cols);
    return l;
}

// Grouped launches of a multi-tile solve: the tiles of one size share every
// launch, tile k owns block k of each working array (its pages, lines or
// row-mask bytes, in tile order) and reads its own inputs through a device
// table of these entries.
struct TileIn {
    const int2* led;     // LED windows (2 x L int32, column-major)
    const void* y;       // image stack (the dtype of the group's YStack)
    uint32_t ymn;        // uint16 stack: raw minimum
    float yrange, yrcp;  // uint16 stack: (max - min) and its reciprocal
    uint32_t pad;
    const float* A;      // pupil mask
    double* loss;        // running loss of the epoch, or null
};

// error plumbing (thread-local last error text)
void set_error(const std::string& msg);
const char* get_error();

struct HipError {
    hipError_t code;
    std::string what;
};

#define FPM_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t _e = (call);                                                            \
        if (_e != hipSuccess)                                                              \
            throw ::fpm::HipError{_e, std::string(#call) + " failed: " + hipGetErrorString(_e) + \
                                          " (" __FILE__ ":" + std::to_string(__LINE__) + ")"}; \
    } while (0)

#define FPM_LAUNCH_CHECK() FPM_HIP(hipGetLastError())

// device twiddle table W_16384^m (FP64-generated), one per device, built on
// first use
const float2* twiddle_table(int device);

}  // namespace fpm
