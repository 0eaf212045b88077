// common.h -- shared host/device definitions of libfpm_amd.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>

namespace fpm {

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
