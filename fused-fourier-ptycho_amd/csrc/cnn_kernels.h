// cnn_kernels.h -- launchers of the warm-start network's kernels (BF16
// matrix-core products with FP32 accumulation, gfx950).
//
// Tensors are NHWC: element (b, y, x, c) of a B x H x W x C tensor sits at
// ((b H + y) W + x) C + c.  Activations and the conv operands are BF16
// (stored as uint16 bit patterns); accumulations, weight gradients, master
// weights and the gradients between layers are FP32.
#pragma once
#include "common.h"

namespace fpm {
namespace cnn {

typedef uint16_t bf16_t;

// 3x3 convolution, zero padding 1, as an implicit GEMM on the matrix cores:
//   out(b, p, co) = bias(co) + sum_{tap, ci} x(b, p + off(tap), ci) w(co, tap, ci)
// w is Cout x 9 x Cin (tap = 3 (dy+1) + (dx+1)); Cin and Cout multiples of 32.
// Epilogue: + bias (if non-null), ReLU (if relu), then stored as BF16 (y) and
// / or FP32 (yf) NHWC.
struct ConvArgs {
    const bf16_t* x;
    const bf16_t* w;
    const float* bias;
    bf16_t* y;
    float* yf;
    int B, H, W, Cin, Cout;
    int relu;
};
void conv3x3(const ConvArgs& a, hipStream_t st);

// C(M x N) += A(M x K) B(N x K)^T, A and B row-major BF16 with K contiguous,
// C row-major FP32, K split over workgroups of ksplit elements each whose
// partial tiles are added with FP32 atomics (C must be cleared first).
// M, N multiples of 32; K any (zero beyond K).
void gemm_nt_atomic(const bf16_t* A, const bf16_t* Bm, float* C, int M, int N, int64_t K, int64_t ksplit,
                    hipStream_t st);

// xcol(tap Cin + ci, q) = x(b, y+dy, x+dx, ci) or 0 off the image, q = (b H + y) W + x;
// a (9 Cin) x Q BF16 matrix with Q = B H W contiguous.
void im2col_t(const bf16_t* x, bf16_t* xcol, int B, int H, int W, int C, hipStream_t st);

// Implicit weight gradient of a 3x3 convolution (no im2col matrix):
//   gW(co, tap, ci) += sum_q gz_t(co, q) xt(ci, q + off(tap))   (0 off the image)
// gz_t: Cout x Q and xt: Cin x Q BF16 (channel-major, Q = B H W contiguous),
// gW FP32 (cleared first; partial sums over pixel ranges added with FP32
// atomics).  Needs W % 8 == 0 (runs of 8 pixels stay in one image row) and
// Cin, Cout multiples of 32.
bool wgrad3x3_supported(int W, int Cin, int Cout);
void wgrad3x3(const bf16_t* gz_t, const bf16_t* xt, float* gW, int B, int H, int W, int Cin, int Cout,
              hipStream_t st);
// channel-major copy of an NHWC tensor: xt(c, q) = x(q, c), Q pixels, C <= 256 channels
void to_channel_major(const bf16_t* x, bf16_t* xt, int64_t Q, int C, hipStream_t st);

// weights for the input gradient: wt(ci, tap, co) = w(co, 8 - tap, ci)
void flip_transpose_w(const bf16_t* w, bf16_t* wt, int Cout, int Cin, hipStream_t st);

// gradient of a conv output before its backward products:
//   gz = gy (times [y > 0] when relu), as BF16 NHWC (for the input gradient)
//   and BF16 C x Q (gz_t, for the weight gradient); db(c) += sum_q gz(q, c)
// (FP32, db must be cleared first).
void grad_prep(const float* gy, const bf16_t* y, int relu, bf16_t* gz, bf16_t* gz_t, float* db, int64_t Q, int C,
               hipStream_t st);

// 2x2 average pooling (H, W even) and its adjoint (gx (+)= gy / 4 broadcast)
void avgpool2_fwd(const bf16_t* x, bf16_t* y, int B, int H, int W, int C, hipStream_t st);
void avgpool2_bwd(const float* gy, float* gx, int B, int H, int W, int C, int accumulate, hipStream_t st);
// nearest 2x upsampling (h, w -> 2h, 2w) and its adjoint (sum of the 4 children)
void upsample2_fwd(const bf16_t* x, bf16_t* y, int B, int h, int w, int C, hipStream_t st);
void upsample2_bwd(const float* gy, float* gx, int B, int h, int w, int C, hipStream_t st);
// y = a + b (BF16 in and out, FP32 sum)
void add_bf16(const bf16_t* a, const bf16_t* b, bf16_t* y, int64_t count, hipStream_t st);
// y (+)= x
void add_f32(const float* x, float* y, int64_t count, int accumulate, hipStream_t st);

// planar FP32 features (B x Cf x H x W) -> BF16 NHWC with C >= Cf channels (zero pad)
void features_to_nhwc(const float* f, bf16_t* x, int B, int Cf, int H, int W, int C, hipStream_t st);

// Head: yf (B x H x W x 2 u^2, FP32) -> planar output out (B x 2 x uH x uW) by
// depth-to-space, channel c u^2 + sy u + sx -> plane c, pixel (u y + sy, u x + sx).
// With target: loss_acc (FP64) += sum (out - target)^2 and
// gyf = 2 (out - target) / count mapped back to the head layout.
void head_shuffle(const float* yf, float* out, int B, int H, int W, int u, hipStream_t st);
void head_mse(const float* yf, const float* target, float* gyf, double* loss_acc, int B, int H, int W, int u,
              double inv_count, hipStream_t st);

// Pupil head (FP32 throughout): per sample b, pooled(c) = mean over the HW
// pixels of e (B x HW x C, BF16, C <= 256), hid = ReLU(W1 pooled + b1)
// (W1: kPupilHidden x C), coef = W2 hid + b2 (W2: K x kPupilHidden, K <= 16).
// ws (B x (C + kPupilHidden + K)) keeps pooled, hid and coef for the
// backward; coef_out (B x K) may be null.
constexpr int kPupilHidden = 64;
constexpr int kPupilMaxTerms = 16;
constexpr int kPupilMaxBatch = 64;
void pupil_head_fwd(const bf16_t* e, int B, int HW, int C, const float* W1, const float* b1, const float* W2,
                    const float* b2, int K, float* ws, float* coef_out, hipStream_t st);
// Backward from target (B x K): with d = coef - target,
// loss_acc (FP64) += lam sum d^2 / (B K); g_coef = 2 lam d / (B K);
// gW2 = g_coef^T hid, gb2 = sum_b g_coef, g_hid = (g_coef W2) [hid > 0],
// gW1 = g_hid^T pooled, gb1 = sum_b g_hid (all written, not accumulated);
// ge (B x HW x C, FP32) += (g_hid W1) / HW at every pixel.  B <= kPupilMaxBatch.
// gpool (B x C) is scratch.
void pupil_head_bwd(const float* ws, const float* target, int B, int HW, int C, const float* W1, const float* W2,
                    int K, float lam, float* gW1, float* gb1, float* gW2, float* gb2, float* gpool, float* ge,
                    double* loss_acc, hipStream_t st);

// Parallax shifts of the brightfield images (the pupil estimate of the warm
// start).  planes: B x P planes of H x W FP32 (row-major, plane (b, p) at
// ((b P + p) H W)).  For sample b and j < count: d(b, j) = (dy, dx), the
// global shift with  img(y, x) ~ ref(y - dy, x - dx), img = plane which[j],
// ref = plane ref, by Lucas-Kanade iterations on the mean/std-normalised
// planes: w(y, x) = img(y + dy, x + dx) (bilinear, coordinates clamped to the
// plane), over the interior pixels g = (central differences of (ref + w) / 2),
// e = w - ref, [sum g g^T] delta = -[sum g e] (FP64 sums), d += delta.
// which: host array of count (<= 64) plane indices; d: device B x count x 2;
// W <= 2730 (a band of rows plus its two halo rows is staged in LDS).
constexpr int kParallaxMaxPlanes = 64;
struct PlaneList {
    int32_t idx[kParallaxMaxPlanes];
};
void parallax_shifts(const float* planes, int B, int P, int H, int W, int ref, const int32_t* which, int count,
                     int iters, float* d, hipStream_t st);

// Adam (no weight decay): m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
// p -= lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps); w16 = bf16(p)
// for the first n16 entries of each layer segment (nullptr: none).
struct AdamArgs {
    float lr, b1, b2, eps, c1, c2;  // c1 = 1/(1-b1^t), c2 = 1/(1-b2^t)
};
void adam_update(float* p, float* m, float* v, const float* g, int64_t count, AdamArgs a, hipStream_t st);
// BF16 copy of FP32 values
void to_bf16(const float* p, bf16_t* w, int64_t count, hipStream_t st);
void fill_f32(float* p, int64_t count, float v, hipStream_t st);

}  // namespace cnn
}  // namespace fpm
