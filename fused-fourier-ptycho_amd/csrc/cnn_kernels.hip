// cnn_kernels.hip -- kernels of the warm-start network on gfx950.
//
// The three products of a 3x3 convolution layer run on the matrix cores as
// v_mfma_f32_16x16x32_bf16 (BF16 operands, FP32 accumulation):
//   forward        out = x * w            implicit GEMM, M = pixels, N = Cout, K = 9 Cin
//   input grad     gx = gz * flip(w)^T    the same kernel on transposed weights
//   weight grad    gw = gz^T im2col(x)    NT GEMM, K = B H W, split over workgroups
// Operand fragments are 8 BF16 (16 bytes) per lane, loaded straight from
// global memory: channel runs of NHWC tensors for the implicit GEMM, rows of
// K-contiguous matrices for the NT GEMM.  Lane layout of the 16x16x32
// product (lane l, i = l mod 16, g = l / 16): A row i, columns 8g..8g+7;
// B column i, rows 8g..8g+7; D column i, rows 4g..4g+3.
#include "cnn_kernels.h"

#include <algorithm>

namespace fpm {
namespace cnn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16_t f2bf(float x) { return __builtin_bit_cast(uint16_t, (__bf16)x); }
__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ bf16x8 ld8(const bf16_t* p) { return __builtin_bit_cast(bf16x8, *(const uint4*)p); }
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0u, 0u, 0u, 0u)); }
__device__ __forceinline__ f32x4 mma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

static unsigned blocks_for(int64_t n, int t) {
    // the element-wise kernels decode indices in 32 bits
    if (n >= ((int64_t)1 << 31)) throw HipError{hipErrorInvalidValue, "cnn: tensor of 2^31 or more elements"};
    return (unsigned)std::min<int64_t>((n + t - 1) / t, 1 << 20);
}

// --------------------------------------------------------------- conv 3x3
// One wave per 16 MT pixels (flattened index within an image) x 32 output
// channels: MT x 2 tiles of 16 x 16, K loop over 9 taps x Cin / 32.  Each
// K step loads MT + 2 fragments for 2 MT products.
template <int MT>
__global__ __launch_bounds__(64) void conv3x3_kernel(ConvArgs a) {
    const int l = threadIdx.x;
    const int i = l & 15, g = l >> 4;
    const int P = a.H * a.W;
    const int p0 = blockIdx.x * 16 * MT;
    const int c0 = blockIdx.y * 32;
    const int b = blockIdx.z;
    const bf16_t* xb = a.x + (size_t)b * P * a.Cin;
    int py[MT], px[MT];
    bool pin[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        const int p = p0 + 16 * mt + i;
        pin[mt] = p < P;
        py[mt] = p / a.W;
        px[mt] = p % a.W;
    }
    f32x4 acc[MT][2];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int tap = 0; tap < 9; ++tap) {
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        const bf16_t* src[MT];
        bool ok[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const int yy = py[mt] + dy, xx = px[mt] + dx;
            ok[mt] = pin[mt] && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W;
            src[mt] = xb + ((size_t)(ok[mt] ? yy : 0) * a.W + (ok[mt] ? xx : 0)) * a.Cin + 8 * g;
        }
        const bf16_t* wr[2];
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) wr[nt] = a.w + ((size_t)(c0 + 16 * nt + i) * 9 + tap) * a.Cin + 8 * g;
        for (int ci = 0; ci < a.Cin; ci += 32) {
            const bf16x8 B0 = ld8(wr[0] + ci);
            const bf16x8 B1 = ld8(wr[1] + ci);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const bf16x8 A = ok[mt] ? ld8(src[mt] + ci) : zero8();
                acc[mt][0] = mma(A, B0, acc[mt][0]);
                acc[mt][1] = mma(A, B1, acc[mt][1]);
            }
        }
    }
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        const int co = c0 + 16 * nt + i;
        const float bias = a.bias ? a.bias[co] : 0.0f;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int p = p0 + 16 * mt + 4 * g + r;
                if (p >= P) continue;
                float v = acc[mt][nt][r] + bias;
                if (a.relu) v = fmaxf(v, 0.0f);
                const size_t o = ((size_t)b * P + p) * a.Cout + co;
                if (a.y) a.y[o] = f2bf(v);
                if (a.yf) a.yf[o] = v;
            }
    }
}

// Patch variant for rows of 64k pixels: a workgroup (4 waves) computes one
// 64-pixel row segment x 32 output channels.  Per chunk of 32 input channels
// it stages the input patch (rows y-1..y+1, pixels x0-1..x0+64, zero off the
// image) and the 32 x 9 x 32 weights in LDS once, and every wave reads the
// fragments of all 9 taps from there (wave w: pixels 16w..16w+15).  A patch
// pixel occupies 80 bytes (32 channels + 8 pad) so that each group of 8
// lanes of a 16-byte LDS read hits distinct banks.
constexpr int kPatchPx = 66, kPatchStride = 40;  // pixels per patch row, BF16 per patch pixel
__global__ __launch_bounds__(256) void conv3x3_patch_kernel(ConvArgs a) {
    __shared__ __align__(16) bf16_t patch[3 * kPatchPx * kPatchStride];
    __shared__ __align__(16) bf16_t wts[32 * 9 * kPatchStride];
    const int t = threadIdx.x, wv = t >> 6, l = t & 63;
    const int i = l & 15, g = l >> 4;
    const int segs = a.W / 64;
    const int y = blockIdx.x / segs, x0 = (blockIdx.x % segs) * 64;
    const int c0 = blockIdx.y * 32, b = blockIdx.z;
    const size_t img = (size_t)b * a.H * a.W;
    f32x4 acc[2] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}};
    for (int ci0 = 0; ci0 < a.Cin; ci0 += 32) {
        __syncthreads();  // the previous chunk's readers are done
        // input patch: 3 x 66 pixels x 4 runs of 8 channels
        for (int e = t; e < 3 * kPatchPx * 4; e += 256) {
            const int run = e & 3, px = (e >> 2) % kPatchPx, r = (e >> 2) / kPatchPx;
            const int yy = y + r - 1, xx = x0 + px - 1;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W)
                v = *(const uint4*)(a.x + (img + (size_t)yy * a.W + xx) * a.Cin + ci0 + 8 * run);
            *(uint4*)(patch + (r * kPatchPx + px) * kPatchStride + 8 * run) = v;
        }
        // weights: 32 output channels x 9 taps x 4 runs of 8 channels
        for (int e = t; e < 32 * 9 * 4; e += 256) {
            const int run = e & 3, ct = e >> 2;  // ct = co * 9 + tap
            const int co = ct / 9, tap = ct % 9;
            *(uint4*)(wts + ct * kPatchStride + 8 * run) =
                *(const uint4*)(a.w + ((size_t)(c0 + co) * 9 + tap) * a.Cin + ci0 + 8 * run);
        }
        __syncthreads();
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3, dx = tap % 3;  // patch row and pixel offset
            const bf16x8 A = __builtin_bit_cast(
                bf16x8, *(const uint4*)(patch + (dy * kPatchPx + 16 * wv + i + dx) * kPatchStride + 8 * g));
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                const bf16x8 B = __builtin_bit_cast(
                    bf16x8, *(const uint4*)(wts + ((16 * nt + i) * 9 + tap) * kPatchStride + 8 * g));
                acc[nt] = mma(A, B, acc[nt]);
            }
        }
    }
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        const int co = c0 + 16 * nt + i;
        const float bias = a.bias ? a.bias[co] : 0.0f;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int x = x0 + 16 * wv + 4 * g + r;
            float v = acc[nt][r] + bias;
            if (a.relu) v = fmaxf(v, 0.0f);
            const size_t o = (img + (size_t)y * a.W + x) * a.Cout + co;
            if (a.y) a.y[o] = f2bf(v);
            if (a.yf) a.yf[o] = v;
        }
    }
}

void conv3x3(const ConvArgs& a, hipStream_t st) {
    if (a.Cin % 32 || a.Cout % 32) throw HipError{hipErrorInvalidValue, "conv3x3: channels must be multiples of 32"};
    if (a.W % 64 == 0) {
        dim3 grid((unsigned)(a.H * (a.W / 64)), a.Cout / 32, a.B);
        hipLaunchKernelGGL(conv3x3_patch_kernel, grid, dim3(256), 0, st, a);
        FPM_LAUNCH_CHECK();
        return;
    }
    const int P = a.H * a.W;
    // 64-pixel waves once there are enough of them to fill the chip, else 32
    const int64_t waves64 = (int64_t)((P + 63) / 64) * (a.Cout / 32) * a.B;
    if (waves64 >= 2048) {
        dim3 grid((P + 63) / 64, a.Cout / 32, a.B);
        hipLaunchKernelGGL(conv3x3_kernel<4>, grid, dim3(64), 0, st, a);
    } else {
        dim3 grid((P + 31) / 32, a.Cout / 32, a.B);
        hipLaunchKernelGGL(conv3x3_kernel<2>, grid, dim3(64), 0, st, a);
    }
    FPM_LAUNCH_CHECK();
}

// ----------------------------------------------------------- NT GEMM
__global__ __launch_bounds__(64) void gemm_nt_kernel(const bf16_t* __restrict__ A, const bf16_t* __restrict__ Bm,
                                                     float* __restrict__ C, int M, int N, int64_t K, int64_t ks) {
    const int l = threadIdx.x;
    const int i = l & 15, g = l >> 4;
    const int m0 = blockIdx.x * 32, n0 = blockIdx.y * 32;
    const int64_t kb = (int64_t)blockIdx.z * ks;
    const int64_t ke = std::min<int64_t>(K, kb + ks);
    const bf16_t* ar[2];
    const bf16_t* br[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        ar[t] = A + (size_t)(m0 + 16 * t + i) * K + 8 * g;
        br[t] = Bm + (size_t)(n0 + 16 * t + i) * K + 8 * g;
    }
    f32x4 acc[2][2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int64_t k = kb; k < ke; k += 32) {
        bf16x8 a0, a1, b0, b1;
        if (k + 32 <= ke) {
            a0 = ld8(ar[0] + k);
            a1 = ld8(ar[1] + k);
            b0 = ld8(br[0] + k);
            b1 = ld8(br[1] + k);
        } else {  // tail: element-wise, zero beyond ke
            a0 = a1 = b0 = b1 = zero8();
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                if (k + 8 * g + e < ke) {
                    a0[e] = __builtin_bit_cast(__bf16, ar[0][k + e]);
                    a1[e] = __builtin_bit_cast(__bf16, ar[1][k + e]);
                    b0[e] = __builtin_bit_cast(__bf16, br[0][k + e]);
                    b1[e] = __builtin_bit_cast(__bf16, br[1][k + e]);
                }
            }
        }
        acc[0][0] = mma(a0, b0, acc[0][0]);
        acc[0][1] = mma(a0, b1, acc[0][1]);
        acc[1][0] = mma(a1, b0, acc[1][0]);
        acc[1][1] = mma(a1, b1, acc[1][1]);
    }
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + 16 * mt + 4 * g + r, col = n0 + 16 * nt + i;
                unsafeAtomicAdd(&C[(size_t)row * N + col], acc[mt][nt][r]);
            }
}

void gemm_nt_atomic(const bf16_t* A, const bf16_t* Bm, float* C, int M, int N, int64_t K, int64_t ksplit,
                    hipStream_t st) {
    if (M % 32 || N % 32 || K % 8 || ksplit % 32 || ksplit <= 0)
        throw HipError{hipErrorInvalidValue, "gemm_nt_atomic: M, N multiples of 32, K of 8, ksplit of 32"};
    dim3 grid(M / 32, N / 32, (unsigned)((K + ksplit - 1) / ksplit));
    hipLaunchKernelGGL(gemm_nt_kernel, grid, dim3(64), 0, st, A, Bm, C, M, N, K, ksplit);
    FPM_LAUNCH_CHECK();
}

// --------------------------------------------- implicit weight gradient
// One wave per (32 output channels, kernel row dy, 32 input channels, pixel
// range): the three taps (dy, -1..+1) of the row share the wave's operands.
// Lane (i, g) holds, per 32-pixel chunk, the run of 8 pixels q .. q+7
// (q = q0 + 8g) of gz_t row co and of xt row ci at row y + dy; the runs for
// dx = -1 / +1 are the aligned run funnel-shifted by one BF16 with its
// neighbour run (zero beyond the row's ends: that is the zero padding).
__device__ __forceinline__ uint4 ld16(const bf16_t* p) { return *(const uint4*)p; }

__global__ __launch_bounds__(64) void wgrad3x3_kernel(const bf16_t* __restrict__ gzt, const bf16_t* __restrict__ xt,
                                                      float* __restrict__ gW, int H, int W, int Cin, int64_t Q,
                                                      int64_t ks) {
    const int l = threadIdx.x;
    const int i = l & 15, g = l >> 4;
    const int co0 = blockIdx.x * 32;
    const int dy = (int)(blockIdx.y % 3) - 1;
    const int ci0 = (int)(blockIdx.y / 3) * 32;
    const int64_t kb = (int64_t)blockIdx.z * ks;
    const int64_t ke = std::min<int64_t>(Q, kb + ks);
    const bf16_t* arow[2];
    const bf16_t* brow[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        arow[t] = gzt + (size_t)(co0 + 16 * t + i) * Q;
        brow[t] = xt + (size_t)(ci0 + 16 * t + i) * Q;
    }
    f32x4 acc[3][2][2];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) acc[d][mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    const uint32_t HWu = (uint32_t)H * (uint32_t)W;
    for (int64_t q0 = kb; q0 < ke; q0 += 32) {
        const int64_t q = q0 + 8 * g;
        bf16x8 a[2];
        uint4 xv[2][3];  // per input-channel tile: runs for dx = -1, 0, +1
        if (q < ke) {
            const uint32_t pix = (uint32_t)(q % (int64_t)HWu);
            const int y = (int)(pix / (uint32_t)W), x0 = (int)(pix % (uint32_t)W);
            const int yy = y + dy;
            const bool rowok = yy >= 0 && yy < H;
            const int64_t src = q + (int64_t)dy * W;  // same image, row y + dy
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                a[t] = ld8(arow[t] + q);
                uint4 vp = make_uint4(0u, 0u, 0u, 0u), v0 = vp, v1 = vp;
                if (rowok) {
                    v0 = ld16(brow[t] + src);
                    if (x0 > 0) vp = ld16(brow[t] + src - 8);
                    if (x0 + 8 < W) v1 = ld16(brow[t] + src + 8);
                }
                xv[t][0] = make_uint4(__builtin_amdgcn_alignbyte(v0.x, vp.w, 2), __builtin_amdgcn_alignbyte(v0.y, v0.x, 2),
                                      __builtin_amdgcn_alignbyte(v0.z, v0.y, 2), __builtin_amdgcn_alignbyte(v0.w, v0.z, 2));
                xv[t][1] = v0;
                xv[t][2] = make_uint4(__builtin_amdgcn_alignbyte(v0.y, v0.x, 2), __builtin_amdgcn_alignbyte(v0.z, v0.y, 2),
                                      __builtin_amdgcn_alignbyte(v0.w, v0.z, 2), __builtin_amdgcn_alignbyte(v1.x, v0.w, 2));
            }
        } else {
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                a[t] = zero8();
                xv[t][0] = xv[t][1] = xv[t][2] = make_uint4(0u, 0u, 0u, 0u);
            }
        }
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                const bf16x8 b = __builtin_bit_cast(bf16x8, xv[nt][d]);
                acc[d][0][nt] = mma(a[0], b, acc[d][0][nt]);
                acc[d][1][nt] = mma(a[1], b, acc[d][1][nt]);
            }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int tap = 3 * (dy + 1) + d;
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int nt = 0; nt < 2; ++nt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int co = co0 + 16 * mt + 4 * g + r, ci = ci0 + 16 * nt + i;
                    unsafeAtomicAdd(&gW[((size_t)co * 9 + tap) * Cin + ci], acc[d][mt][nt][r]);
                }
    }
}

bool wgrad3x3_supported(int W, int Cin, int Cout) { return W % 8 == 0 && Cin % 32 == 0 && Cout % 32 == 0; }

void wgrad3x3(const bf16_t* gz_t, const bf16_t* xt, float* gW, int B, int H, int W, int Cin, int Cout,
              hipStream_t st) {
    if (!wgrad3x3_supported(W, Cin, Cout) || B < 1 || H < 1)
        throw HipError{hipErrorInvalidValue, "wgrad3x3: W a multiple of 8, channels multiples of 32"};
    const int64_t Q = (int64_t)B * H * W;
    // pixel ranges (multiples of 32 pixels, at least 512) so that a launch has
    // about 2048 waves: the 32 x 32 layers have only 3 tiles to spread
    const int64_t tiles = (int64_t)(Cout / 32) * 3 * (Cin / 32);
    const int64_t splits = std::max<int64_t>(1, std::min<int64_t>((2048 + tiles - 1) / tiles, Q / 512));
    const int64_t ks = std::max<int64_t>(512, ((Q + splits - 1) / splits + 31) / 32 * 32);
    dim3 grid(Cout / 32, 3 * (Cin / 32), (unsigned)((Q + ks - 1) / ks));
    hipLaunchKernelGGL(wgrad3x3_kernel, grid, dim3(64), 0, st, gz_t, xt, gW, H, W, Cin, Q, ks);
    FPM_LAUNCH_CHECK();
}

// channel-major copy through an LDS tile of 64 pixels x C channels
__global__ __launch_bounds__(256) void to_channel_major_kernel(const bf16_t* __restrict__ x, bf16_t* __restrict__ xt,
                                                               int64_t Q, int C) {
    __shared__ bf16_t tile[64 * 257];
    const int t = threadIdx.x;
    const int elems = 64 * C;
    for (int64_t q0 = (int64_t)blockIdx.x * 64; q0 < Q; q0 += (int64_t)gridDim.x * 64) {
        __syncthreads();
        for (int e = t; e < elems; e += 256) {
            const int qq = e / C, c = e - qq * C;
            tile[qq * 257 + c] = q0 + qq < Q ? x[(size_t)(q0 + qq) * C + c] : (bf16_t)0;
        }
        __syncthreads();
        for (int e = t; e < elems; e += 256) {
            const int c = e / 64, qq = e % 64;
            if (q0 + qq < Q) xt[(size_t)c * Q + q0 + qq] = tile[qq * 257 + c];
        }
    }
}

void to_channel_major(const bf16_t* x, bf16_t* xt, int64_t Q, int C, hipStream_t st) {
    if (C < 1 || C > 256) throw HipError{hipErrorInvalidValue, "to_channel_major: 1 <= C <= 256"};
    if (Q <= 0) return;
    const unsigned blocks = (unsigned)std::min<int64_t>((Q + 63) / 64, 4096);
    hipLaunchKernelGGL(to_channel_major_kernel, dim3(blocks), dim3(256), 0, st, x, xt, Q, C);
    FPM_LAUNCH_CHECK();
}

// ------------------------------------------------------------ im2col^T
// thread: one pixel q and 8 channels; writes 9 x 8 rows of xcol at column q
__global__ __launch_bounds__(256) void im2col_t_kernel(const bf16_t* __restrict__ x, bf16_t* __restrict__ xcol, int B,
                                                       int H, int W, int C) {
    const int64_t Q = (int64_t)B * H * W;
    const int c8 = C / 8;
    const int64_t total = Q * c8;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
        // 32-bit index arithmetic (the launcher checks the sizes): 64-bit
        // division is a long instruction sequence on the GPU
        const uint32_t q = (uint32_t)t % (uint32_t)Q;
        const uint32_t cg = (uint32_t)t / (uint32_t)Q;
        const uint32_t xx = q % (uint32_t)W;
        const uint32_t qw = q / (uint32_t)W;
        const uint32_t yy = qw % (uint32_t)H;
        const uint32_t b = qw / (uint32_t)H;
        for (int tap = 0; tap < 9; ++tap) {
            const int sy = (int)yy + tap / 3 - 1, sx = (int)xx + tap % 3 - 1;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (sy >= 0 && sy < H && sx >= 0 && sx < W)
                v = *(const uint4*)(x + (((size_t)b * H + sy) * W + sx) * C + 8 * cg);
            const uint16_t* e = (const uint16_t*)&v;
#pragma unroll
            for (int k = 0; k < 8; ++k) xcol[((size_t)tap * C + 8 * cg + k) * Q + q] = e[k];
        }
    }
}

void im2col_t(const bf16_t* x, bf16_t* xcol, int B, int H, int W, int C, hipStream_t st) {
    if (C % 8) throw HipError{hipErrorInvalidValue, "im2col_t: channels must be a multiple of 8"};
    const int64_t total = (int64_t)B * H * W * (C / 8);
    hipLaunchKernelGGL(im2col_t_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, st, x, xcol, B, H, W, C);
    FPM_LAUNCH_CHECK();
}

__global__ void flip_transpose_w_kernel(const bf16_t* __restrict__ w, bf16_t* __restrict__ wt, int Cout, int Cin) {
    const int64_t total = (int64_t)Cout * 9 * Cin;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
        const uint32_t u = (uint32_t)t, uc = u / (uint32_t)Cout;
        const int co = (int)(u % (uint32_t)Cout);
        const int tap = (int)(uc % 9u);
        const int ci = (int)(uc / 9u);
        wt[t] = w[((size_t)co * 9 + (8 - tap)) * Cin + ci];  // wt(ci, tap, co)
    }
}

void flip_transpose_w(const bf16_t* w, bf16_t* wt, int Cout, int Cin, hipStream_t st) {
    const int64_t total = (int64_t)Cout * 9 * Cin;
    hipLaunchKernelGGL(flip_transpose_w_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, st, w, wt, Cout, Cin);
    FPM_LAUNCH_CHECK();
}

// ---------------------------------------------------------- grad prep
// workgroup: chunks of 64 consecutive q (grid-stride), all C (<= 256)
// channels through an LDS tile; per-channel bias sums stay in LDS across the
// chunks and end in one atomic per channel per workgroup
constexpr int kGpQ = 64;
constexpr unsigned kGpMaxBlocks = 1024;
__global__ __launch_bounds__(256) void grad_prep_kernel(const float* __restrict__ gy, const bf16_t* __restrict__ y,
                                                        int relu, bf16_t* __restrict__ gz, bf16_t* __restrict__ gz_t,
                                                        float* __restrict__ db, int64_t Q, int C) {
    __shared__ bf16_t tile[kGpQ * 257];
    __shared__ float dbs[256];
    const int t = threadIdx.x;
    const int elems = kGpQ * C;  // <= 16384
    for (int c = t; c < C; c += 256) dbs[c] = 0.0f;
    for (int64_t q0 = (int64_t)blockIdx.x * kGpQ; q0 < Q; q0 += (int64_t)gridDim.x * kGpQ) {
        __syncthreads();  // the previous chunk's transposed reads are done
        for (int e = t; e < elems; e += 256) {
            const int qq = e / C, c = e - qq * C;
            const int64_t q = q0 + qq;
            bf16_t h = 0;
            if (q < Q) {
                const size_t o = (size_t)q * C + c;
                float v = gy[o];
                if (relu && !(bf2f(y[o]) > 0.0f)) v = 0.0f;
                h = f2bf(v);
                gz[o] = h;
            }
            tile[qq * 257 + c] = h;
        }
        __syncthreads();
        // per-channel sums of the rounded values (thread c owns channel c)
        for (int c = t; c < C; c += 256) {
            float s = 0.0f;
            for (int qq = 0; qq < kGpQ; ++qq) s += bf2f(tile[qq * 257 + c]);
            dbs[c] += s;
        }
        for (int e = t; e < elems; e += 256) {
            const int c = e / kGpQ, qq = e % kGpQ;
            const int64_t q = q0 + qq;
            if (q < Q) gz_t[(size_t)c * Q + q] = tile[qq * 257 + c];
        }
    }
    for (int c = t; c < C; c += 256) unsafeAtomicAdd(&db[c], dbs[c]);
}

void grad_prep(const float* gy, const bf16_t* y, int relu, bf16_t* gz, bf16_t* gz_t, float* db, int64_t Q, int C,
               hipStream_t st) {
    if (C > 256) throw HipError{hipErrorInvalidValue, "grad_prep: at most 256 channels"};
    if (Q <= 0) return;
    const unsigned blocks = (unsigned)std::min<int64_t>((Q + kGpQ - 1) / kGpQ, kGpMaxBlocks);
    hipLaunchKernelGGL(grad_prep_kernel, dim3(blocks), dim3(256), 0, st, gy, y, relu, gz, gz_t, db, Q, C);
    FPM_LAUNCH_CHECK();
}

// ------------------------------------------------- pooling / upsampling
__global__ void avgpool2_fwd_kernel(const bf16_t* __restrict__ x, bf16_t* __restrict__ y, int B, int H, int W, int C) {
    const int h = H / 2, w = W / 2;
    const int64_t total = (int64_t)B * h * w * C;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
        const uint32_t u = (uint32_t)t, u1 = u / (uint32_t)C, u2 = u1 / (uint32_t)w;
        const int c = (int)(u % (uint32_t)C);
        const int xx = (int)(u1 % (uint32_t)w);
        const int yy = (int)(u2 % (uint32_t)h);
        const int64_t b = (int64_t)(u2 / (uint32_t)h);
        const size_t r0 = ((b * H + 2 * yy) * W + 2 * xx) * C + c;
        const size_t r1 = r0 + (size_t)W * C;
        const float s = (bf2f(x[r0]) + bf2f(x[r0 + C])) + (bf2f(x[r1]) + bf2f(x[r1 + C]));
        y[t] = f2bf(s * 0.25f);
    }
}

__global__ void avgpool2_bwd_kernel(const float* __restrict__ gy, float* __restrict__ gx, int B, int H, int W, int C,
                                    int acc) {
    const int64_t total = (int64_t)B * H * W * C;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
        const uint32_t u = (uint32_t)t, u1 = u / (uint32_t)C, u2 = u1 / (uint32_t)W;
        const int c = (int)(u % (uint32_t)C);
        const int xx = (int)(u1 % (uint32_t)W);
        const int yy = (int)(u2 % (uint32_t)H);
        const int64_t b = (int64_t)(u2 / (uint32_t)H);
        const float v = gy[((b * (H / 2) + yy / 2) * (W / 2) + xx / 2) * C + c] * 0.25f;
        gx[t] = acc ? gx[t] + v : v;
    }
}

__global__ void upsample2_fwd_kernel(const bf16_t* __restrict__ x, bf16_t* __restrict__ y, int B, int h, int w,
                                     int C) {
    const int H = 2 * h, W = 2 * w;
    const int64_t total = (int64_t)B * H * W * C;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
        const uint32_t u = (uint32_t)t, u1 = u / (uint32_t)C, u2 = u1 / (uint32_t)W;
        const int c = (int)(u % (uint32_t)C);
        const int xx = (int)(u1 % (uint32_t)W);
        const int yy = (int)(u2 % (uint32_t)H);
        const int64_t b = (int64_t)(u2 / (uint32_t)H);
        y[t] = x[((b * h + yy / 2) * w + xx / 2) * C + c];
    }
}

__global__ void upsample2_bwd_kernel(const float* __restrict__ gy, float* __restrict__ gx, int B, int h, int w, int C) {
    const int W = 2 * w;
    const int64_t total = (int64_t)B * h * w * C;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
        const uint32_t u = (uint32_t)t, u1 = u / (uint32_t)C, u2 = u1 / (uint32_t)w;
        const int c = (int)(u % (uint32_t)C);
        const int xx = (int)(u1 % (uint32_t)w);
        const int yy = (int)(u2 % (uint32_t)h);
        const int64_t b = (int64_t)(u2 / (uint32_t)h);
        const size_t r0 = ((b * 2 * h + 2 * yy) * W + 2 * xx) * C + c;
        const size_t r1 = r0 + (size_t)W * C;
        gx[t] = (gy[r0] + gy[r0 + C]) + (gy[r1] + gy[r1 + C]);
    }
}

void avgpool2_fwd(const bf16_t* x, bf16_t* y, int B, int H, int W, int C, hipStream_t st) {
    const int64_t total = (int64_t)B * (H / 2) * (W / 2) * C;
    hipLaunchKernelGGL(avgpool2_fwd_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, st, x, y, B, H, W, C);
    FPM_LAUNCH_CHECK();
}
void avgpool2_bwd(const float* gy, float* gx, int B, int H, int W, int C, int accumulate, hipStream_t st) {
    const int64_t total = (int64_t)B * H * W * C;
    hipLaunchKernelGGL(avgpool2_bwd_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, st, gy, gx, B, H, W, C,
                       accumulate);
    FPM_LAUNCH_CHECK();
}
void upsample2_fwd(const bf16_t* x, bf16_t* y, int B, int h, int w, int C, hipStream_t st) {
    const int64_t total = (int64_t)B * 4 * h * w * C;
    hipLaunchKernelGGL(upsample2_fwd_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, st, x, y, B, h, w, C);
    FPM_LAUNCH_CHECK();
}
void upsample2_bwd(const float* gy, float* gx, int B, int h, int w, int C, hipStream_t st) {
    const int64_t total = (int64_t)B * h * w * C;
    hipLaunchKernelGGL(upsample2_bwd_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, st, gy, gx, B, h, w, C);
    FPM_LAUNCH_CHECK();
}

// ---------------------------------------------------------- elementwise
__global__ void add_bf16_kernel(const bf16_t* __restrict__ a, const bf16_t* __restrict__ b, bf16_t* __restrict__ y,
                                int64_t n) {
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256)
        y[t] = f2bf(bf2f(a[t]) + bf2f(b[t]));
}
__global__ void add_f32_kernel(const float* __restrict__ x, float* __restrict__ y, int64_t n, int acc) {
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256)
        y[t] = acc ? y[t] + x[t] : x[t];
}
void add_bf16(const bf16_t* a, const bf16_t* b, bf16_t* y, int64_t count, hipStream_t st) {
    hipLaunchKernelGGL(add_bf16_kernel, dim3(blocks_for(count, 256)), dim3(256), 0, st, a, b, y, count);
    FPM_LAUNCH_CHECK();
}
void add_f32(const float* x, float* y, int64_t count, int accumulate, hipStream_t st) {
    hipLaunchKernelGGL(add_f32_kernel, dim3(blocks_for(count, 256)), dim3(256), 0, st, x, y, count, accumulate);
    FPM_LAUNCH_CHECK();
}

__global__ void features_kernel(const float* __restrict__ f, bf16_t* __restrict__ x, int B, int Cf, int H, int W,
                                int C) {
    const int64_t total = (int64_t)B * H * W * C;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
        const uint32_t u = (uint32_t)t, u1 = u / (uint32_t)C;
        const int c = (int)(u % (uint32_t)C);
        const int64_t p = (int64_t)(u1 % (uint32_t)(H * W));
        const int64_t b = (int64_t)(u1 / (uint32_t)(H * W));
        x[t] = c < Cf ? f2bf(f[(b * Cf + c) * H * W + p]) : (bf16_t)0;
    }
}
void features_to_nhwc(const float* f, bf16_t* x, int B, int Cf, int H, int W, int C, hipStream_t st) {
    const int64_t total = (int64_t)B * H * W * C;
    hipLaunchKernelGGL(features_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, st, f, x, B, Cf, H, W, C);
    FPM_LAUNCH_CHECK();
}

// -------------------------------------------------------------- head
// head element e = ((b H + y) W + x) 2u^2 + ch, ch = c u^2 + sy u + sx
__device__ __forceinline__ size_t head_out_index(int64_t e, int H, int W, int u) {
    // 32-bit index arithmetic (the launchers check the sizes)
    const uint32_t U2 = (uint32_t)(u * u), CH = 2u * U2;
    const uint32_t ev = (uint32_t)e, pix = ev / CH, pw = pix / (uint32_t)W;
    const uint32_t ch = ev % CH;
    const uint32_t x = pix % (uint32_t)W;
    const uint32_t y = pw % (uint32_t)H;
    const uint32_t b = pw / (uint32_t)H;
    const uint32_t c = ch / U2, sy = (ch % U2) / (uint32_t)u, sx = ch % (uint32_t)u;
    const size_t OH = (size_t)u * H, OW = (size_t)u * W;
    return (((size_t)b * 2 + c) * OH + (size_t)u * y + sy) * OW + (size_t)u * x + sx;
}

__global__ void head_shuffle_kernel(const float* __restrict__ yf, float* __restrict__ out, int B, int H, int W, int u) {
    const int64_t total = (int64_t)B * H * W * 2 * u * u;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256)
        out[head_out_index(e, H, W, u)] = yf[e];
}

__global__ __launch_bounds__(256) void head_mse_kernel(const float* __restrict__ yf, const float* __restrict__ target,
                                                       float* __restrict__ gyf, double* __restrict__ loss_acc, int B,
                                                       int H, int W, int u, double inv_count) {
    __shared__ double red[4];
    const int64_t total = (int64_t)B * H * W * 2 * u * u;
    double s = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
        const float d = yf[e] - target[head_out_index(e, H, W, u)];
        s += (double)d * (double)d;
        gyf[e] = (float)(2.0 * (double)d * inv_count);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0 && loss_acc) unsafeAtomicAdd(loss_acc, red[0] + red[1] + red[2] + red[3]);
}

void head_shuffle(const float* yf, float* out, int B, int H, int W, int u, hipStream_t st) {
    const int64_t total = (int64_t)B * H * W * 2 * u * u;
    hipLaunchKernelGGL(head_shuffle_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, st, yf, out, B, H, W, u);
    FPM_LAUNCH_CHECK();
}
void head_mse(const float* yf, const float* target, float* gyf, double* loss_acc, int B, int H, int W, int u,
              double inv_count, hipStream_t st) {
    const int64_t total = (int64_t)B * H * W * 2 * u * u;
    // every workgroup ends in one FP64 atomic on the same address: few
    // workgroups, each striding over many elements
    const unsigned blocks = std::min(blocks_for(total, 256), 1024u);
    hipLaunchKernelGGL(head_mse_kernel, dim3(blocks), dim3(256), 0, st, yf, target, gyf, loss_acc, B, H, W, u,
                       inv_count);
    FPM_LAUNCH_CHECK();
}

// --------------------------------------------------------- pupil head
// one workgroup per sample; thread c sums channel c over the pixels
// (consecutive threads read consecutive channels of a pixel)
__global__ __launch_bounds__(256) void pupil_head_fwd_kernel(const bf16_t* __restrict__ e, int HW, int C,
                                                             const float* __restrict__ W1, const float* __restrict__ b1,
                                                             const float* __restrict__ W2, const float* __restrict__ b2,
                                                             int K, float* __restrict__ ws,
                                                             float* __restrict__ coef_out) {
    __shared__ float pooled[256];
    __shared__ float hid[kPupilHidden];
    const int t = threadIdx.x, b = blockIdx.x;
    const int stride = C + kPupilHidden + K;
    float* w = ws + (size_t)b * stride;
    if (t < C) {
        const bf16_t* eb = e + (size_t)b * HW * C + t;
        float s = 0.0f;
        for (int p = 0; p < HW; ++p) s += bf2f(eb[(size_t)p * C]);
        pooled[t] = s / (float)HW;
        w[t] = pooled[t];
    }
    __syncthreads();
    if (t < kPupilHidden) {
        float s = b1[t];
        for (int c = 0; c < C; ++c) s = fmaf(W1[t * C + c], pooled[c], s);
        hid[t] = fmaxf(s, 0.0f);
        w[C + t] = hid[t];
    }
    __syncthreads();
    if (t < K) {
        float s = b2[t];
        for (int j = 0; j < kPupilHidden; ++j) s = fmaf(W2[t * kPupilHidden + j], hid[j], s);
        w[C + kPupilHidden + t] = s;
        if (coef_out) coef_out[b * K + t] = s;
    }
}

// one workgroup for the whole batch: every parameter gradient is a sum over
// the B samples in a fixed order (deterministic)
__global__ __launch_bounds__(256) void pupil_head_bwd_kernel(const float* __restrict__ ws,
                                                             const float* __restrict__ target, int B, int HW, int C,
                                                             const float* __restrict__ W1, const float* __restrict__ W2,
                                                             int K, float lam, float* __restrict__ gW1,
                                                             float* __restrict__ gb1, float* __restrict__ gW2,
                                                             float* __restrict__ gb2, float* __restrict__ gpool,
                                                             double* __restrict__ loss_acc) {
    __shared__ float gc[kPupilMaxBatch * kPupilMaxTerms];
    __shared__ float gh[kPupilMaxBatch * kPupilHidden];
    const int t = threadIdx.x;
    const int stride = C + kPupilHidden + K;
    const float inv = 1.0f / (float)(B * K);
    for (int i = t; i < B * K; i += 256) {
        const int b = i / K, k = i - b * K;
        gc[i] = ws[(size_t)b * stride + C + kPupilHidden + k] - target[i];  // d, scaled below
    }
    __syncthreads();
    if (t == 0 && loss_acc) {
        double s = 0.0;
        for (int i = 0; i < B * K; ++i) s += (double)gc[i] * (double)gc[i];
        *loss_acc += (double)lam * s / (double)(B * K);
    }
    __syncthreads();
    for (int i = t; i < B * K; i += 256) gc[i] *= 2.0f * lam * inv;
    __syncthreads();
    for (int i = t; i < B * kPupilHidden; i += 256) {
        const int b = i / kPupilHidden, j = i - b * kPupilHidden;
        float s = 0.0f;
        if (ws[(size_t)b * stride + C + j] > 0.0f)
            for (int k = 0; k < K; ++k) s = fmaf(gc[b * K + k], W2[k * kPupilHidden + j], s);
        gh[i] = s;
    }
    __syncthreads();
    for (int i = t; i < K * kPupilHidden; i += 256) {
        const int k = i / kPupilHidden, j = i - k * kPupilHidden;
        float s = 0.0f;
        for (int b = 0; b < B; ++b) s = fmaf(gc[b * K + k], ws[(size_t)b * stride + C + j], s);
        gW2[i] = s;
    }
    for (int k = t; k < K; k += 256) {
        float s = 0.0f;
        for (int b = 0; b < B; ++b) s += gc[b * K + k];
        gb2[k] = s;
    }
    for (int i = t; i < kPupilHidden * C; i += 256) {
        const int j = i / C, c = i - j * C;
        float s = 0.0f;
        for (int b = 0; b < B; ++b) s = fmaf(gh[b * kPupilHidden + j], ws[(size_t)b * stride + c], s);
        gW1[i] = s;
    }
    for (int j = t; j < kPupilHidden; j += 256) {
        float s = 0.0f;
        for (int b = 0; b < B; ++b) s += gh[b * kPupilHidden + j];
        gb1[j] = s;
    }
    for (int i = t; i < B * C; i += 256) {
        const int b = i / C, c = i - b * C;
        float s = 0.0f;
        for (int j = 0; j < kPupilHidden; ++j) s = fmaf(gh[b * kPupilHidden + j], W1[j * C + c], s);
        gpool[i] = s / (float)HW;
    }
}

// ge(b, p, c) += gpool(b, c)
__global__ void pupil_pool_bwd_kernel(const float* __restrict__ gpool, float* __restrict__ ge, int B, int HW, int C) {
    const uint32_t total = (uint32_t)B * (uint32_t)HW * (uint32_t)C;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < total; i += gridDim.x * 256u) {
        const uint32_t c = i % (uint32_t)C, b = i / ((uint32_t)HW * (uint32_t)C);
        ge[i] += gpool[b * (uint32_t)C + c];
    }
}

void pupil_head_fwd(const bf16_t* e, int B, int HW, int C, const float* W1, const float* b1, const float* W2,
                    const float* b2, int K, float* ws, float* coef_out, hipStream_t st) {
    if (C > 256 || K < 1 || K > kPupilMaxTerms || B < 1 || HW < 1)
        throw HipError{hipErrorInvalidValue, "pupil_head_fwd: C <= 256, 1 <= K <= 16, B, HW >= 1"};
    hipLaunchKernelGGL(pupil_head_fwd_kernel, dim3(B), dim3(256), 0, st, e, HW, C, W1, b1, W2, b2, K, ws, coef_out);
    FPM_LAUNCH_CHECK();
}

void pupil_head_bwd(const float* ws, const float* target, int B, int HW, int C, const float* W1, const float* W2,
                    int K, float lam, float* gW1, float* gb1, float* gW2, float* gb2, float* gpool, float* ge,
                    double* loss_acc, hipStream_t st) {
    if (C > 256 || K < 1 || K > kPupilMaxTerms || B < 1 || B > kPupilMaxBatch || HW < 1)
        throw HipError{hipErrorInvalidValue, "pupil_head_bwd: C <= 256, 1 <= K <= 16, 1 <= B <= 64, HW >= 1"};
    hipLaunchKernelGGL(pupil_head_bwd_kernel, dim3(1), dim3(256), 0, st, ws, target, B, HW, C, W1, W2, K, lam, gW1,
                       gb1, gW2, gb2, gpool, loss_acc);
    FPM_LAUNCH_CHECK();
    const int64_t total = (int64_t)B * HW * C;
    hipLaunchKernelGGL(pupil_pool_bwd_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, st, gpool, ge, B, HW, C);
    FPM_LAUNCH_CHECK();
}

// ------------------------------------------------------ parallax shifts
constexpr int kPxThreads = 512;
constexpr int kPxLdsFloats = 8192;  // per staged plane (32 KiB): (band + 2) x W values

__device__ __forceinline__ double block_sum_px(double v, double* red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    __syncthreads();  // red is free (previous readers are done)
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kPxThreads / 64; ++k) s += red[k];
    return s;
}

// bilinear sample of plane q (H, W >= 2) at (y, x), coordinates clamped to
// the plane: the cell's top-left corner is (min(floor y, H-2), min(floor x, W-2))
__device__ __forceinline__ float bilinear_clamped(const float* __restrict__ q, int H, int W, float y, float x) {
    y = fminf(fmaxf(y, 0.0f), (float)(H - 1));
    x = fminf(fmaxf(x, 0.0f), (float)(W - 1));
    const int y0 = min((int)y, H - 2), x0 = min((int)x, W - 2);
    const float fy = y - (float)y0, fx = x - (float)x0;
    const float* r0 = q + (size_t)y0 * W + x0;
    const float* r1 = r0 + W;
    const float a = fmaf(fx, r0[1] - r0[0], r0[0]);
    const float b = fmaf(fx, r1[1] - r1[0], r1[0]);
    return fmaf(fy, b - a, a);
}

// One workgroup per (image j, sample b).  Each iteration walks the interior
// rows in bands: the normalised reference rows and the warped, normalised
// image rows of a band (plus one halo row above and below) are staged in
// LDS, so every warped value is sampled once and its neighbours come from LDS.
__global__ __launch_bounds__(kPxThreads) void parallax_shift_kernel(const float* __restrict__ planes, int P, int H,
                                                                    int W, int ref, PlaneList which, int count,
                                                                    int iters, float* __restrict__ dout) {
    __shared__ double red[kPxThreads / 64];
    __shared__ float rs[kPxLdsFloats];
    __shared__ float ws[kPxLdsFloats];
    const int j = blockIdx.x, b = blockIdx.y, t = threadIdx.x;
    const size_t HW = (size_t)H * W;
    const float* R = planes + ((size_t)b * P + ref) * HW;
    const float* I = planes + ((size_t)b * P + which.idx[j]) * HW;
    // mean and standard deviation of both planes (FP64)
    double sr = 0.0, sr2 = 0.0, si = 0.0, si2 = 0.0;
    for (size_t k = t; k < HW; k += kPxThreads) {
        const double r = R[k], v = I[k];
        sr += r;
        sr2 += r * r;
        si += v;
        si2 += v * v;
    }
    sr = block_sum_px(sr, red);
    sr2 = block_sum_px(sr2, red);
    si = block_sum_px(si, red);
    si2 = block_sum_px(si2, red);
    const double inv = 1.0 / (double)HW;
    const double mr = sr * inv, mi = si * inv;
    const double vr = sr2 * inv - mr * mr, vi = si2 * inv - mi * mi;
    const float ar = vr > 0.0 ? (float)(1.0 / sqrt(vr)) : 0.0f, ai = vi > 0.0 ? (float)(1.0 / sqrt(vi)) : 0.0f;
    const float br = (float)mr, bi = (float)mi;
    const int band = kPxLdsFloats / W - 2;  // interior rows per band (>= 1: W <= kPxLdsFloats / 3)
    const int iw = W - 2;
    double dy = 0.0, dx = 0.0;
    for (int it = 0; it < iters; ++it) {
        const float fy = (float)dy, fx = (float)dx;
        double s00 = 0.0, s01 = 0.0, s11 = 0.0, b0 = 0.0, b1 = 0.0;
        for (int y0 = 1; y0 < H - 1; y0 += band) {
            const int rows = min(band, H - 1 - y0);  // interior rows y0 .. y0 + rows - 1
            const int staged = (rows + 2) * W;        // rows y0 - 1 .. y0 + rows
            __syncthreads();                          // the previous band's readers are done
            for (int k = t; k < staged; k += kPxThreads) {
                const int yy = y0 - 1 + k / W, xx = k % W;
                rs[k] = (R[(size_t)yy * W + xx] - br) * ar;
                ws[k] = (bilinear_clamped(I, H, W, (float)yy + fy, (float)xx + fx) - bi) * ai;
            }
            __syncthreads();
            for (int k = t; k < rows * iw; k += kPxThreads) {
                const int ly = 1 + k / iw, x = 1 + k % iw;  // LDS row of the pixel, column
                const int c = ly * W + x;
                const double g0 = 0.25 * ((double)(rs[c + W] - rs[c - W]) + (double)(ws[c + W] - ws[c - W]));
                const double g1 = 0.25 * ((double)(rs[c + 1] - rs[c - 1]) + (double)(ws[c + 1] - ws[c - 1]));
                const double e = (double)ws[c] - (double)rs[c];
                s00 += g0 * g0;
                s01 += g0 * g1;
                s11 += g1 * g1;
                b0 += g0 * e;
                b1 += g1 * e;
            }
        }
        s00 = block_sum_px(s00, red);
        s01 = block_sum_px(s01, red);
        s11 = block_sum_px(s11, red);
        b0 = block_sum_px(b0, red);
        b1 = block_sum_px(b1, red);
        const double det = s00 * s11 - s01 * s01;
        if (!(det > 0.0)) break;  // no texture: keep the current estimate (every thread sees the same sums)
        dy -= (s11 * b0 - s01 * b1) / det;
        dx -= (s00 * b1 - s01 * b0) / det;
    }
    if (t == 0) {
        dout[((size_t)b * count + j) * 2 + 0] = (float)dy;
        dout[((size_t)b * count + j) * 2 + 1] = (float)dx;
    }
}

void parallax_shifts(const float* planes, int B, int P, int H, int W, int ref, const int32_t* which, int count,
                     int iters, float* d, hipStream_t st) {
    if (B < 1 || B > 65535 || P < 1 || H < 3 || W < 3 || W > kPxLdsFloats / 3 || ref < 0 || ref >= P || count < 1 ||
        count > kParallaxMaxPlanes || iters < 0 || (int64_t)H * W >= ((int64_t)1 << 31))
        throw HipError{hipErrorInvalidValue,
                       "parallax_shifts: B in [1, 65535], planes of at least 3 x 3 with at most 2730 columns, "
                       "0 <= ref < P, 1 <= count <= 64, iters >= 0"};
    PlaneList pl{};
    for (int j = 0; j < count; ++j) {
        if (which[j] < 0 || which[j] >= P) throw HipError{hipErrorInvalidValue, "parallax_shifts: plane index >= P"};
        pl.idx[j] = which[j];
    }
    hipLaunchKernelGGL(parallax_shift_kernel, dim3(count, B), dim3(kPxThreads), 0, st, planes, P, H, W, ref, pl, count,
                       iters, d);
    FPM_LAUNCH_CHECK();
}

// -------------------------------------------------------------- Adam
__global__ void adam_kernel(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                            const float* __restrict__ g, int64_t n, AdamArgs a) {
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
        const float gr = g[t];
        const float mm = a.b1 * m[t] + (1.0f - a.b1) * gr;
        const float vv = a.b2 * v[t] + (1.0f - a.b2) * gr * gr;
        m[t] = mm;
        v[t] = vv;
        p[t] -= a.lr * (mm * a.c1) / (sqrtf(vv * a.c2) + a.eps);
    }
}
void adam_update(float* p, float* m, float* v, const float* g, int64_t count, AdamArgs a, hipStream_t st) {
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(count, 256)), dim3(256), 0, st, p, m, v, g, count, a);
    FPM_LAUNCH_CHECK();
}

__global__ void to_bf16_kernel(const float* __restrict__ p, bf16_t* __restrict__ w, int64_t n) {
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) w[t] = f2bf(p[t]);
}
void to_bf16(const float* p, bf16_t* w, int64_t count, hipStream_t st) {
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(count, 256)), dim3(256), 0, st, p, w, count);
    FPM_LAUNCH_CHECK();
}

__global__ void fill_kernel(float* __restrict__ p, int64_t n, float v) {
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) p[t] = v;
}
void fill_f32(float* p, int64_t count, float v, hipStream_t st) {
    hipLaunchKernelGGL(fill_kernel, dim3(blocks_for(count, 256)), dim3(256), 0, st, p, count, v);
    FPM_LAUNCH_CHECK();
}

}  // namespace cnn
}  // namespace fpm
