// optim.hip -- the reference's MATLAB-side optimizers and pupil constraint as
// device kernels, for solves driven by the gradient-only entry
// (fpm_fused_gradient): the loop of codes/main_run_FPM.m:106-160 and
// benchmark/cuFPM_recon_2025_circular.m:152-183 runs
//   [dldw1, dldw2] = fullyfusedFPM(w1, w2, Y, ledIdx, pratio);
//   w1 = optimizer_w1.step(w1, dldw1);  w2 = optimizer_w2.step(w2, dldw2);
//   w2 = max(min(abs(w2), P0 + ac), P0 - ac) .* sign(w2) .* P0;
// with one of
//   optimizer_adam     codes/func_ware/optimizers/optimizer_adam.m:25-38
//   optimizer_yogi     codes_v2/func_ware/optimizers/optimizer_yogi.m:25-36
//   optimizer_RMSprop  codes/func_ware/optimizers/optimizer_RMSprop.m:28-46
//                      (as the callers construct it: centered false, mum 0)
//   optimizer_sgd      codes/func_ware/optimizers/optimizer_sgd.m:16-21
// MATLAB evaluates these on single gpuArrays: a double scalar meets a single
// array as a single, and a double-only subexpression such as (1 - beta) is
// rounded to single once -- the host does those scalar steps in FP64 and the
// kernel gets the rounded constants, so each element sees the same float
// operations as MATLAB's.  abs(g).^2 is |g| squared (MATLAB's abs of a
// complex single is its modulus, here sqrt(x x + y y) in single).  Floating
// contraction is off in these kernels: each product and sum rounds on its
// own, as in the MATLAB expressions.
#include "../../include/fpm_amd.h"
#include "common.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>

namespace fpm {
void set_error(const std::string& msg);

namespace {

struct OptConst {
    int32_t kind;
    float lr;     // (float)lr
    float b1;     // (float)beta1
    float c1;     // (float)(1 - beta1)
    float b2;     // (float)beta2
    float c2;     // (float)(1 - beta2)
    float bias;   // Adam: (float)(sqrt(1 - beta2^iter) / (1 - beta1^iter))
};

__device__ __forceinline__ float msign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

__global__ __launch_bounds__(256) void optimizer_step_kernel(float2* __restrict__ p, const float2* __restrict__ g,
                                                             float2* __restrict__ m1, float* __restrict__ m2,
                                                             size_t n, OptConst c) {
#pragma clang fp contract(off)
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256) {
        const float2 gr = g[k];
        float2 w = p[k];
        if (c.kind == FPM_OPT_SGD) {
            w.x -= c.lr * gr.x;
            w.y -= c.lr * gr.y;
            p[k] = w;
            continue;
        }
        const float ag = sqrtf(gr.x * gr.x + gr.y * gr.y);
        const float g2 = ag * ag;
        if (c.kind == FPM_OPT_RMSPROP) {
            // squave = r1 + (1 - r1) |g|^2: no running average (optimizer_RMSprop.m:31)
            const float sq = c.b1 + c.c1 * g2;
            const float den = sqrtf(sq) + 1e-5f;
            w.x -= c.lr * (gr.x / den);
            w.y -= c.lr * (gr.y / den);
            p[k] = w;
            continue;
        }
        float2 a = m1[k];
        a.x = c.b1 * a.x + c.c1 * gr.x;
        a.y = c.b1 * a.y + c.c1 * gr.y;
        float v = m2[k];
        if (c.kind == FPM_OPT_ADAM) {
            v = c.b2 * v + c.c2 * g2;
        } else {  // Yogi: v - (1 - r2) sign(v - |g|^2) |g|^2, MATLAB sign(0) = 0
            v = v - c.c2 * msign(v - g2) * g2;
        }
        const float den = sqrtf(v) + 1e-5f;
        if (c.kind == FPM_OPT_ADAM) {
            // update = bias .* mom1 ./ (sqrt(mom2) + 1e-5), left to right
            w.x -= c.lr * ((c.bias * a.x) / den);
            w.y -= c.lr * ((c.bias * a.y) / den);
        } else {
            w.x -= c.lr * (a.x / den);
            w.y -= c.lr * (a.y / den);
        }
        m1[k] = a;
        m2[k] = v;
        p[k] = w;
    }
}

// w2 = max(min(|w2|, A + ac), A - ac) .* sign(w2) .* A, sign(w) = w / |w|
// (0 at w = 0), in MATLAB's order: clamp, times sign, times A
__global__ __launch_bounds__(256) void pupil_constraint_kernel(float2* __restrict__ P, const float* __restrict__ A,
                                                               size_t n, float ac) {
#pragma clang fp contract(off)
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256) {
        const float2 w = P[k];
        const float a = A[k];
        const float m = sqrtf(w.x * w.x + w.y * w.y);
        const float c = fmaxf(fminf(m, a + ac), a - ac);
        float2 s = make_float2(0.f, 0.f);
        if (m > 0.0f) s = make_float2(w.x / m, w.y / m);
        P[k] = make_float2(c * s.x * a, c * s.y * a);
    }
}

unsigned blocks_for(size_t n) { return (unsigned)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 1u << 16)); }

int guarded(const char* fn, const std::function<void()>& body) {
    set_error("");
    try {
        body();
        return FPM_OK;
    } catch (const HipError& e) {
        set_error(std::string(fn) + ": " + e.what);
        return FPM_E_HIP;
    } catch (const std::invalid_argument& e) {
        set_error(std::string(fn) + ": " + e.what());
        return FPM_E_INVALID;
    }
}

}  // namespace
}  // namespace fpm

using namespace fpm;

extern "C" {

int fpm_optimizer_step(int32_t kind, fpm_c32* param, const fpm_c32* grad, fpm_c32* mom1, float* mom2, int64_t count,
                       double lr, double beta1, double beta2, int64_t iter, void* stream) {
    return guarded("fpm_optimizer_step", [&] {
        if (kind != FPM_OPT_SGD && kind != FPM_OPT_ADAM && kind != FPM_OPT_YOGI && kind != FPM_OPT_RMSPROP)
            throw std::invalid_argument("kind must be FPM_OPT_SGD, FPM_OPT_ADAM, FPM_OPT_YOGI or FPM_OPT_RMSPROP");
        if (!param || !grad) throw std::invalid_argument("null array argument");
        if ((kind == FPM_OPT_ADAM || kind == FPM_OPT_YOGI) && (!mom1 || !mom2))
            throw std::invalid_argument("Adam and Yogi need both moment arrays");
        if (count < 0) throw std::invalid_argument("count must be >= 0");
        if (kind == FPM_OPT_ADAM && iter < 1) throw std::invalid_argument("Adam: iter must be >= 1");
        OptConst c{};
        c.kind = kind;
        c.lr = (float)lr;
        c.b1 = (float)beta1;
        c.c1 = (float)(1.0 - beta1);
        c.b2 = (float)beta2;
        c.c2 = (float)(1.0 - beta2);
        if (kind == FPM_OPT_ADAM) {
            const double den = 1.0 - std::pow(beta1, (double)iter);
            if (den == 0.0) throw std::invalid_argument("Adam: 1 - beta1^iter is zero");
            c.bias = (float)(std::sqrt(1.0 - std::pow(beta2, (double)iter)) / den);
        }
        if (count == 0) return;
        hipLaunchKernelGGL(optimizer_step_kernel, dim3(blocks_for((size_t)count)), dim3(256), 0, (hipStream_t)stream,
                           (float2*)param, (const float2*)grad, (float2*)mom1, mom2, (size_t)count, c);
        FPM_LAUNCH_CHECK();
    });
}

int fpm_pupil_constraint(fpm_c32* P, const float* A, int64_t count, double amplitude_constraint, void* stream) {
    return guarded("fpm_pupil_constraint", [&] {
        if (!P || !A) throw std::invalid_argument("null array argument");
        if (count < 0) throw std::invalid_argument("count must be >= 0");
        if (count == 0) return;
        hipLaunchKernelGGL(pupil_constraint_kernel, dim3(blocks_for((size_t)count)), dim3(256), 0, (hipStream_t)stream,
                           (float2*)P, A, (size_t)count, (float)amplitude_constraint);
        FPM_LAUNCH_CHECK();
    });
}

}  // extern "C"
