// generic.hip -- the solve for even sides that are not powers of two (up to
// 16384: field-of-view tiles of 192 px -> 768 px, 200 px -> 1600 px, 1536 px
// -> 12288 px, and the like -- any even side, as cuFFT takes any length),
// which the power-of-two kernels cannot take (their index maps are bit masks
// and their FFT passes radix-16 slot maps).  The steps are those of
// cuFPM_v3/cuFPM_pure_v3.cu:128-207 without the fusions, in natural order:
//   S = DFT_N(O); crop x pupil (kernels.cuh:19-65); z = IDFT_n; loss stencil
//   (kernels.cuh:197-275); R = DFT_n(u); owner-computes stitch and the pupil
//   gradient / RMSprop + clamp (kernels.cuh:67-129, 384-410);
//   G_O = IDFT_N(G_S); RMSprop of O (kernels.cuh:364-382).
// The transforms are a self-sorting (Stockham) mixed-radix FFT: radices 4, 2,
// 3, 5 and 7 with hard-coded butterflies, any other prime factor as a direct
// sum; pass (r, n, s) maps
//   y[q + s (r p + k)] = (sum_j x[q + s (p + j n/r)] W_r^(j k)) W_n^(p k),
// lines staged in LDS (each pass in place through registers), twiddles from
// an FP64-generated table of W_len^k.
#include "../../include/fpm_amd.h"
#include "common.h"
#include "fpm_kernels.h"
#include "generic.h"
#include "fft_core.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

#ifndef FPM_GEN_CONTIG_POINTS
#define FPM_GEN_CONTIG_POINTS 1024
#endif
#ifndef FPM_GEN_STRIDED_POINTS
#define FPM_GEN_STRIDED_POINTS 2048
#endif
#ifndef FPM_GEN_SET_ATTR
#define FPM_GEN_SET_ATTR 1
#endif
#ifndef FPM_GEN_LB
#define FPM_GEN_LB 1024
#endif
// crop in the first row pass and the fused column steps (gen_cols_fused_kernel)
#ifndef FPM_GEN_FUSED
#define FPM_GEN_FUSED 1
#endif
#ifndef FPM_GEN_FUSED_MAX
#define FPM_GEN_FUSED_MAX 2048
#endif
// lines of 3 2^k points as a radix-3 step + three power-of-two transforms (fft3m_kernel)
#ifndef FPM_GEN_3M
#define FPM_GEN_3M 1
#endif
// object update fused with the contiguous passes for N_rows = 3 2^k (gen_update3m_kernel)
#ifndef FPM_GEN_FUSED_UPDATE
#define FPM_GEN_FUSED_UPDATE 1
#endif
// launch shapes of the transform passes measured at a solve's set-up
#ifndef FPM_GEN_TUNE
#define FPM_GEN_TUNE 1
#endif

namespace fpm {

bool generic_side(int64_t v) { return v >= 16 && v <= kGenericMaxSide && (v & 1) == 0; }

namespace {

// v mod n: one subtraction for v < 2n, which every index sum below stays
// under for LED offsets from fpm_led_offsets (-n/2 + 1 <= k <= N - n/2); a
// caller's offsets outside that range take the division and stay in bounds
__device__ __forceinline__ uint32_t wrap1(uint32_t v, uint32_t n) {
    return v < n ? v : (v - n < n ? v - n : v % n);
}

__device__ __forceinline__ float2 cmul_g(float2 a, float2 b) {  // the roundings of fpm::cmul
    return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cmulc_g(float2 a, float2 b) {  // a conj(b), the roundings of fpm::cmulc
    return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.y, b.x, -(a.x * b.y)));
}

// x / d for a divisor fixed per launch: q = (umulhi(x, m) + x) >> s with
// s = ceil(log2 d), m = floor(2^32 (2^s - d) / d) + 1 (Granlund and
// Montgomery's round-up method; exact for every 32-bit x)
struct FastDiv {
    uint32_t d, m, s;
};
FastDiv make_fastdiv(uint32_t d) {
    FastDiv f{d, 0, 0};
    while ((1ull << f.s) < d) ++f.s;
    f.m = (uint32_t)(((1ull << 32) * ((1ull << f.s) - d)) / d + 1);
    return f;
}
__device__ __forceinline__ uint32_t fdiv(uint32_t x, const FastDiv& f) {
    return (uint32_t)(((uint64_t)__umulhi(x, f.m) + x) >> f.s);
}

constexpr int kGenMaxFactors = 32;
// points of a line one thread holds in a pass at most: the workgroup has at
// least (lines x length) / kGenPts threads
constexpr int kGenPts = 16;

struct GenFFT {
    const float2* in;
    float2* out;
    const float2* tw;  // W_len^k, k < len
    uint64_t batch_stride, line_stride, es;
    uint32_t len, nlines, lines_per_batch, tl, nf, inverse;
    FastDiv f_lpb, f_tl, f_len;
    FastDiv f_nb[kGenMaxFactors];  // len / R of each pass
    FastDiv f_s[kGenMaxFactors];   // the stride s of each pass
    FastDiv f_r[kGenMaxFactors];   // R of each pass
    uint16_t fac[kGenMaxFactors];
};

// element e of line L at (L / lpb) batch_stride + (L % lpb) line_stride + e es
__device__ __forceinline__ size_t gen_off(const GenFFT& a, uint32_t L, uint32_t e) {
    const uint32_t pg = fdiv(L, a.f_lpb);
    return (size_t)pg * a.batch_stride + (size_t)(L - pg * a.lines_per_batch) * a.line_stride + (size_t)e * a.es;
}

// -i v (forward) or +i v (inverse)
__device__ __forceinline__ float2 mul_mi(float2 v, bool inv) {
    return inv ? make_float2(-v.y, v.x) : make_float2(v.y, -v.x);
}

// radix-R DFT of v in place (forward: W_R = exp(-2 pi i / R); inverse: conjugate)
template <int R> __device__ __forceinline__ void small_dft(float2* v, bool inv);
template <> __device__ __forceinline__ void small_dft<2>(float2* v, bool) {
    const float2 a = v[0], b = v[1];
    v[0] = make_float2(a.x + b.x, a.y + b.y);
    v[1] = make_float2(a.x - b.x, a.y - b.y);
}
template <> __device__ __forceinline__ void small_dft<4>(float2* v, bool inv) {
    const float2 s02 = make_float2(v[0].x + v[2].x, v[0].y + v[2].y);
    const float2 d02 = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
    const float2 s13 = make_float2(v[1].x + v[3].x, v[1].y + v[3].y);
    const float2 d13 = make_float2(v[1].x - v[3].x, v[1].y - v[3].y);
    // forward: X1 = d02 - i d13, X3 = d02 + i d13 (inverse swaps them)
    const float2 mi = mul_mi(d13, inv);
    v[0] = make_float2(s02.x + s13.x, s02.y + s13.y);
    v[2] = make_float2(s02.x - s13.x, s02.y - s13.y);
    v[1] = make_float2(d02.x + mi.x, d02.y + mi.y);
    v[3] = make_float2(d02.x - mi.x, d02.y - mi.y);
}
template <> __device__ __forceinline__ void small_dft<3>(float2* v, bool inv) {
    // X0 = a + b + c; X1,2 = a - (b + c)/2 -+ i (sqrt(3)/2)(b - c) (forward), +- for the inverse
    const float c3 = 0.86602540378443864676f;
    const float2 t = make_float2(v[1].x + v[2].x, v[1].y + v[2].y);
    const float2 d = make_float2(v[1].x - v[2].x, v[1].y - v[2].y);
    const float2 m = make_float2(v[0].x - 0.5f * t.x, v[0].y - 0.5f * t.y);
    const float2 r = mul_mi(make_float2(c3 * d.x, c3 * d.y), inv);
    v[0] = make_float2(v[0].x + t.x, v[0].y + t.y);
    v[1] = make_float2(m.x + r.x, m.y + r.y);
    v[2] = make_float2(m.x - r.x, m.y - r.y);
}
// odd R: pair x_j with x_{R-j}: t_j = x_j + x_{R-j}, d_j = x_j - x_{R-j};
// X_k = x_0 + sum_j cos(2 pi j k / R) t_j -+ i sum_j sin(2 pi j k / R) d_j,
// X_{R-k} the same with the sign of the i term flipped (forward: X_k takes -)
template <int R>
__device__ __forceinline__ void odd_dft(float2* v, bool inv, const float* c, const float* sn) {
    constexpr int H = R / 2;
    float2 t[H], d[H];
#pragma unroll
    for (int j = 1; j <= H; ++j) {
        t[j - 1] = make_float2(v[j].x + v[R - j].x, v[j].y + v[R - j].y);
        d[j - 1] = make_float2(v[j].x - v[R - j].x, v[j].y - v[R - j].y);
    }
    float2 x0 = v[0];
    float2 s0 = x0;
#pragma unroll
    for (int j = 0; j < H; ++j) s0 = make_float2(s0.x + t[j].x, s0.y + t[j].y);
#pragma unroll
    for (int k = 1; k <= H; ++k) {
        float2 ak = x0, bk = make_float2(0.f, 0.f);
#pragma unroll
        for (int j = 1; j <= H; ++j) {
            const int m = (j * k) % R;  // cos(m th) = c[min(m, R-m)], sin(m th) = +-sn[min(m, R-m)]
            const int mm = m <= H ? m : R - m;
            const float cs = c[mm - 1], ss = m <= H ? sn[mm - 1] : -sn[mm - 1];
            ak = make_float2(ak.x + cs * t[j - 1].x, ak.y + cs * t[j - 1].y);
            bk = make_float2(bk.x + ss * d[j - 1].x, bk.y + ss * d[j - 1].y);
        }
        const float2 ib = mul_mi(bk, inv);
        v[k] = make_float2(ak.x + ib.x, ak.y + ib.y);
        v[R - k] = make_float2(ak.x - ib.x, ak.y - ib.y);
    }
    v[0] = s0;
}
template <> __device__ __forceinline__ void small_dft<5>(float2* v, bool inv) {
    const float c[2] = {0.30901699437494742410f, -0.80901699437494742410f};  // cos(2 pi k / 5)
    const float sn[2] = {0.95105651629515357212f, 0.58778525229247312917f};  // sin(2 pi k / 5)
    odd_dft<5>(v, inv, c, sn);
}
template <> __device__ __forceinline__ void small_dft<7>(float2* v, bool inv) {
    const float c[3] = {0.62348980185873353053f, -0.22252093395631440429f, -0.90096886790241912624f};
    const float sn[3] = {0.78183148246802980871f, 0.97492791218182360702f, 0.43388373911755812048f};
    odd_dft<7>(v, inv, c, sn);
}

// One in-place pass (R, n, s) of a line set held in LDS x:
//   x'[q + s (R p + k)] = (sum_j x[q + s (p + j n/R)] W_R^(j k)) W_n^(p k)
// every thread computes its butterflies into registers, the workgroup
// synchronises, then writes them back (one buffer instead of two, so a
// 16384-point line fits the LDS)
template <int R>
__device__ __forceinline__ void gen_pass(const GenFFT& a, float2* x, uint32_t nl, uint32_t n, uint32_t s,
                                         const FastDiv& fs, const FastDiv& fnb, bool inv) {
    constexpr int IT = (kGenPts + R - 1) / R;
    const uint32_t m = n / R, nb = a.len / R, step_n = a.len / n, items = nl * nb;
    float2 res[IT][R];
    uint32_t dst[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const uint32_t bi = threadIdx.x + (uint32_t)it * blockDim.x;
        dst[it] = 0xffffffffu;
        if (bi < items) {
            const uint32_t l = fdiv(bi, fnb), b = bi - l * nb;
            const uint32_t p = fdiv(b, fs), q = b - p * s;
            const float2* xl = x + (size_t)l * a.len;
            float2 v[R];
#pragma unroll
            for (int j = 0; j < R; ++j) v[j] = xl[q + s * (p + j * m)];
            small_dft<R>(v, inv);
            res[it][0] = v[0];
#pragma unroll
            for (int k = 1; k < R; ++k) {
                float2 w = a.tw[p * (uint32_t)k * step_n];  // p k < (n / R) R = n
                if (inv) w.y = -w.y;
                res[it][k] = cmul_g(v[k], w);
            }
            dst[it] = l * a.len + q + s * (R * p);
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it)
        if (dst[it] != 0xffffffffu) {
#pragma unroll
            for (int k = 0; k < R; ++k) x[dst[it] + s * (uint32_t)k] = res[it][k];
        }
    __syncthreads();
}

// A pass of any other (prime) radix R, one output per work item: X_k of
// butterfly (p, q) is a direct R-point sum, W_R^(j k) = W_len^((j k mod R)
// len / R) from the same table (R len / R products per line and pass)
__device__ __forceinline__ void gen_pass_any(const GenFFT& a, float2* x, uint32_t nl, uint32_t n, uint32_t s, uint32_t R,
                                          const FastDiv& fs, const FastDiv& fr, bool inv) {
    const uint32_t m = n / R, step_n = a.len / n, step_r = a.len / R, items = nl * a.len;
    float2 res[kGenPts];
    uint32_t dst[kGenPts];
#pragma unroll
    for (int it = 0; it < kGenPts; ++it) {
        const uint32_t idx = threadIdx.x + (uint32_t)it * blockDim.x;
        dst[it] = 0xffffffffu;
        if (idx < items) {
            const uint32_t l = fdiv(idx, a.f_len), e = idx - l * a.len;
            const uint32_t b = fdiv(e, fr), k = e - b * R;  // butterfly b, output k
            const uint32_t p = fdiv(b, fs), q = b - p * s;
            const float2* xl = x + (size_t)l * a.len + q + s * p;
            float2 acc = make_float2(0.f, 0.f);
            uint32_t jk = 0;
            for (uint32_t j = 0; j < R; ++j) {
                float2 w = a.tw[jk * step_r];
                if (inv) w.y = -w.y;
                const float2 t = cmul_g(xl[s * j * m], w);
                acc = make_float2(acc.x + t.x, acc.y + t.y);
                jk += k;
                jk = jk >= R ? jk - R : jk;
            }
            float2 w = a.tw[p * k * step_n];
            if (inv) w.y = -w.y;
            res[it] = cmul_g(acc, w);
            dst[it] = l * a.len + q + s * (R * p + k);
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kGenPts; ++it)
        if (dst[it] != 0xffffffffu) x[dst[it]] = res[it];
    __syncthreads();
}

// every pass of the transform of the nl lines in x (LDS), in place
__device__ __forceinline__ void gen_passes(const GenFFT& a, float2* x, uint32_t nl, bool inv) {
    uint32_t n = a.len, s = 1;
    for (uint32_t f = 0; f < a.nf; ++f) {
        const uint32_t r = a.fac[f];
        switch (r) {
            case 4: gen_pass<4>(a, x, nl, n, s, a.f_s[f], a.f_nb[f], inv); break;
            case 2: gen_pass<2>(a, x, nl, n, s, a.f_s[f], a.f_nb[f], inv); break;
            case 3: gen_pass<3>(a, x, nl, n, s, a.f_s[f], a.f_nb[f], inv); break;
            case 5: gen_pass<5>(a, x, nl, n, s, a.f_s[f], a.f_nb[f], inv); break;
            case 7: gen_pass<7>(a, x, nl, n, s, a.f_s[f], a.f_nb[f], inv); break;
            default: gen_pass_any(a, x, nl, n, s, r, a.f_s[f], a.f_r[f], inv); break;
        }
        n /= r;
        s *= r;
    }
}

// The crop x pupil as the source of a row pass (kernels.cuh:38-64): line L
// is row di of page pg = L / n_r of the batch (pages of tile k = pg / b at
// b k .. b k + b - 1 with tab), element dj its column; the value is
// ifftshift(S P) of LED led0 + pg mod b, formed as gen_crop_kernel forms it.
struct GenCrop {
    const float2* S;
    const float2* P;
    const int2* led;
    const TileIn* tab;
    uint32_t led0, b, nr, nc, Nr, Nc;
    FastDiv f_b;
};
// the P and S elements behind gen_crop_value (S is read only where P != 0)
struct GenCropAt {
    const float2* p;
    const float2* s;
};
__device__ __forceinline__ GenCropAt gen_crop_at(const GenCrop& c, uint32_t pg, uint32_t di, uint32_t dj) {
    const float2* S = c.S;
    const float2* P = c.P;
    const int2* led = c.led;
    uint32_t l = pg;
    if (c.tab) {
        const uint32_t tk = fdiv(pg, c.f_b);
        l = pg - tk * c.b;
        S += (size_t)tk * c.Nr * c.Nc;
        P += (size_t)tk * c.nr * c.nc;
        led = c.tab[tk].led;
    }
    const uint32_t i = wrap1(di + c.nr / 2, c.nr), j = wrap1(dj + c.nc / 2, c.nc);
    const int2 kk = led[c.led0 + l];
    const uint32_t srow = wrap1(i + (uint32_t)kk.y - 1u + c.Nr / 2, c.Nr);
    const uint32_t scol = wrap1(j + (uint32_t)kk.x - 1u + c.Nc / 2, c.Nc);
    return GenCropAt{P + (size_t)j * c.nr + i, S + (size_t)scol * c.Nr + srow};
}
__device__ __forceinline__ float2 gen_crop_value(const GenCrop& c, uint32_t pg, uint32_t di, uint32_t dj) {
    const float2* S = c.S;
    const float2* P = c.P;
    const int2* led = c.led;
    uint32_t l = pg;
    if (c.tab) {
        const uint32_t tk = fdiv(pg, c.f_b);
        l = pg - tk * c.b;
        S += (size_t)tk * c.Nr * c.Nc;
        P += (size_t)tk * c.nr * c.nc;
        led = c.tab[tk].led;
    }
    const uint32_t i = wrap1(di + c.nr / 2, c.nr), j = wrap1(dj + c.nc / 2, c.nc);
    const int2 kk = led[c.led0 + l];
    const uint32_t srow = wrap1(i + (uint32_t)kk.y - 1u + c.Nr / 2, c.Nr);
    const uint32_t scol = wrap1(j + (uint32_t)kk.x - 1u + c.Nc / 2, c.Nc);
    const float2 pp = P[(size_t)j * c.nr + i];
    float2 z = make_float2(0.f, 0.f);
    if (pp.x != 0.0f || pp.y != 0.0f) z = cmul_g(S[(size_t)scol * c.Nr + srow], pp);
    return z;
}

template <bool CROP>
__global__ __launch_bounds__(FPM_GEN_LB) void fft_generic_kernel(GenFFT a, GenCrop c) {
    extern __shared__ float2 sm[];
    float2* x = sm;
    const uint32_t L0 = blockIdx.x * a.tl;
    const uint32_t nl = min(a.tl, a.nlines - L0);
    const uint32_t tot = nl * a.len;
    const bool line_fast = a.es != 1;  // strided lines: adjacent lines are adjacent in memory
    // (line, element) of a load/store index: line fastest for strided lines
    auto split = [&](uint32_t idx) -> uint2 {
        if (line_fast) {
            const uint32_t e = nl == a.tl ? fdiv(idx, a.f_tl) : idx / nl;
            return make_uint2(idx - e * nl, e);
        }
        const uint32_t l = fdiv(idx, a.f_len);
        return make_uint2(l, idx - l * a.len);
    };
    for (uint32_t idx = threadIdx.x; idx < tot; idx += blockDim.x) {
        const uint2 le = split(idx);
        float2 v;
        if constexpr (CROP) {
            const uint32_t L = L0 + le.x, pg = fdiv(L, a.f_lpb);
            v = gen_crop_value(c, pg, L - pg * a.lines_per_batch, le.y);
        } else {
            v = a.in[gen_off(a, L0 + le.x, le.y)];
        }
        x[(size_t)le.x * a.len + le.y] = v;
    }
    __syncthreads();
    gen_passes(a, x, nl, a.inverse != 0);
    for (uint32_t idx = threadIdx.x; idx < tot; idx += blockDim.x) {
        const uint2 le = split(idx);
        a.out[gen_off(a, L0 + le.x, le.y)] = x[(size_t)le.x * a.len + le.y];
    }
}

// ---- lines of N = 3 M points (M = 2^k, 16 <= M <= 4096): 192, 384, 768, 1536 ...
// Cooley-Tukey with N = 3 x M, n = n2 + M n1, k = k1 + 3 k2:
//   X[k1 + 3 k2] = sum_n2 W_M^(n2 k2) (W_N^(n2 k1) sum_n1 x[n2 + M n1] W_3^(n1 k1))
// A workgroup stages NL = P / M lines in LDS (natural order, coalesced from
// global memory whichever way the lines lie), forms the radix-3 step
// (small_dft<3>'s arithmetic) and the twiddle W_N^(n2 k1) (the FP64-made
// table of W_N) straight into the first-pass registers of 3 NL sub-lines of
// M points, runs the power-of-two Stockham passes of fft_core.h on them
// (radix 16 in registers, one LDS exchange per pass instead of the generic
// engine's one per radix-4/2/3 factor) and stages X[k1 + 3 k2] back in
// natural order for a coalesced store.  3 P / 16 threads, 3 P (17/16) words
// of LDS: P = 1024 for contiguous lines (192 threads, 26 KiB: several
// workgroups per CU overlap their load, transform and store phases), P =
// 16 M for strided lines (16 adjacent lines make 128-byte runs).
// The transform of nl <= NL lines of N = 3 M points staged in LDS (natural
// order, pitch N + 1), in place: radix-3 step + twiddle into the first-pass
// registers of the 3 NL sub-lines, the power-of-two passes, X[k1 + 3 k2]
// back to the staged layout.  twN: W_N^k, k < N.  Ends with a barrier.
template <int M, int NL> struct Lines3m {
    static constexpr int N = 3 * M, SUB = 3 * NL, T = SUB * M / 16, LD = N + 1;
    using Map = LineSlowMap<SUB, M>;
    static constexpr int WORDS = Map::LDS_WORDS > NL * LD ? Map::LDS_WORDS : NL * LD;
};
template <int M, int NL, bool INV>
__device__ __forceinline__ void fft3m_lines(float2* lds, uint32_t nl, const float2* __restrict__ twN,
                                            const Twiddles& tw, int t) {
    using C = Lines3m<M, NL>;
    constexpr int T = C::T, LD = C::LD;
    using SL = Slots<M, INV, T, typename C::Map, ENG_F32>;
    float2 v[16];
#pragma unroll
    for (int q = 0; q < SL::NG0; ++q)
#pragma unroll
        for (int s = 0; s < SL::GS0; ++s) {
            int line, n2;
            SL::in(t, q, s, line, n2);
            const uint32_t l = (uint32_t)line / 3u, k1 = (uint32_t)line - 3u * l;
            float2 y = make_float2(0.f, 0.f);
            if (l < nl) {
                const float2* xl = lds + l * LD + n2;
                float2 x3[3] = {xl[0], xl[M], xl[2 * M]};
                small_dft<3>(x3, INV);
                y = k1 == 0 ? x3[0] : (k1 == 1 ? x3[1] : x3[2]);
                if (k1 != 0) {
                    float2 w = twN[(uint32_t)n2 * k1];  // n2 k1 < 2 M < N
                    if (INV) w.y = -w.y;
                    y = cmul_g(y, w);
                }
            }
            v[q * SL::GS0 + s] = y;
        }
    first_pass_dft<M, INV, T>(v);
    __syncthreads();
    remaining_passes<M, INV, T, typename C::Map>(v, lds, t, tw);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < SL::NGL; ++q)
#pragma unroll
        for (int s = 0; s < SL::GSL; ++s) {
            int line, k2;
            SL::out(t, q, s, line, k2);
            const uint32_t l = (uint32_t)line / 3u, k1 = (uint32_t)line - 3u * l;
            if (l < nl) lds[l * LD + k1 + 3u * (uint32_t)k2] = v[q * SL::GSL + s];
        }
    __syncthreads();
}

template <int M, int P, bool INV, bool CROP>
__global__ __launch_bounds__(3 * P / 16) void fft3m_kernel(GenFFT a, GenCrop c, Twiddles tw) {
    constexpr int N = 3 * M, NL = P / M, T = 3 * P / 16, LD = N + 1;
    static_assert(NL >= 1 && NL * N == 16 * T, "16 points per thread");
    __shared__ float2 lds[Lines3m<M, NL>::WORDS];
    __shared__ float2 twl[2 * kTwLo];
    const int t = threadIdx.x;
    tw = bind_twiddles(tw, twl, t, T);
    const uint32_t L0 = blockIdx.x * (uint32_t)NL;
    const uint32_t nl = min((uint32_t)NL, a.nlines - L0);
    const uint32_t tot = nl * (uint32_t)N;
    const bool line_fast = a.es != 1;  // strided lines: adjacent lines adjacent in memory
    // (line, element) of a staging index: line fastest for strided lines
    auto split = [&](uint32_t idx, uint32_t& l, uint32_t& e) {
        if (line_fast) {
            e = nl == (uint32_t)NL ? idx / (uint32_t)NL : idx / nl;
            l = idx - e * nl;
        } else {
            l = idx / (uint32_t)N;
            e = idx - l * (uint32_t)N;
        }
    };
    auto load = [&](uint32_t l, uint32_t e) -> float2 {
        if constexpr (CROP) {
            const uint32_t L = L0 + l, pg = fdiv(L, a.f_lpb);
            return gen_crop_value(c, pg, L - pg * a.lines_per_batch, e);
        } else {
            return a.in[gen_off(a, L0 + l, e)];
        }
    };
    // a full workgroup moves exactly 16 points per thread: unrolled, so every
    // thread keeps its 16 loads in flight at once
    if (nl == (uint32_t)NL) {
        float2 buf[16];
        if constexpr (CROP) {
            // the crop's P loads, then its S loads where P != 0, then the
            // products (gen_crop_value's operations): two round trips
            float2 pv[16];
            auto at = [&](int i) {
                uint32_t l, e;
                split((uint32_t)(t + i * T), l, e);
                const uint32_t L = L0 + l, pg = fdiv(L, a.f_lpb);
                return gen_crop_at(c, pg, L - pg * a.lines_per_batch, e);
            };
#pragma unroll
            for (int i = 0; i < 16; ++i) pv[i] = *at(i).p;
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                buf[i] = make_float2(0.f, 0.f);
                if (pv[i].x != 0.0f || pv[i].y != 0.0f) buf[i] = *at(i).s;
            }
#pragma unroll
            for (int i = 0; i < 16; ++i)
                buf[i] = (pv[i].x != 0.0f || pv[i].y != 0.0f) ? cmul_g(buf[i], pv[i]) : make_float2(0.f, 0.f);
        } else {
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                uint32_t l, e;
                split((uint32_t)(t + i * T), l, e);
                buf[i] = load(l, e);
            }
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            uint32_t l, e;
            split((uint32_t)(t + i * T), l, e);
            lds[l * LD + e] = buf[i];
        }
    } else {
        for (uint32_t idx = t; idx < tot; idx += T) {
            uint32_t l, e;
            split(idx, l, e);
            lds[l * LD + e] = load(l, e);
        }
    }
    __syncthreads();
    fft3m_lines<M, NL, INV>(lds, nl, a.tw, tw, t);
    if (nl == (uint32_t)NL) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            uint32_t l, e;
            split((uint32_t)(t + i * T), l, e);
            a.out[gen_off(a, L0 + l, e)] = lds[l * LD + e];
        }
    } else {
        for (uint32_t idx = t; idx < tot; idx += T) {
            uint32_t l, e;
            split(idx, l, e);
            a.out[gen_off(a, L0 + l, e)] = lds[l * LD + e];
        }
    }
}

// M of a length that the 3 x 2^k kernel takes (0: none); strided lines only
// up to 768 points (16 lines of a workgroup make 128-byte runs)
uint32_t len3m(uint32_t len, bool strided) {
    if (!FPM_GEN_3M || len % 3 != 0) return 0;
    const uint32_t m = len / 3;
    if (m < 16 || m > 4096 || (m & (m - 1)) != 0) return 0;
    if (strided && m > 256) return 0;
    return m;
}

template <int M, int P, bool INV, bool CROP>
void launch_3m_shape(const GenFFT& a, const GenCrop& c, const Twiddles& tw, hipStream_t st) {
    constexpr int NL = P / M;
    hipLaunchKernelGGL((fft3m_kernel<M, P, INV, CROP>), dim3((a.nlines + NL - 1) / NL), dim3(3 * P / 16), 0, st, a, c,
                       tw);
}

template <bool INV, bool CROP>
void launch_3m(uint32_t m, const GenFFT& a, const GenCrop& c, hipStream_t st) {
    int dev = 0;
    FPM_HIP(hipGetDevice(&dev));
    const Twiddles tw{twiddle_table(dev)};
    const bool strided = a.es != 1;
#define FPM_3M_CASE(MM, PC, PS)                                                    \
    case MM:                                                                       \
        if (strided)                                                               \
            launch_3m_shape<MM, PS, INV, CROP>(a, c, tw, st);                      \
        else                                                                       \
            launch_3m_shape<MM, PC, INV, CROP>(a, c, tw, st);                      \
        break;
    switch (m) {
        FPM_3M_CASE(16, 1024, 1024)
        FPM_3M_CASE(32, 1024, 1024)
        FPM_3M_CASE(64, 1024, 1024)
        FPM_3M_CASE(128, 1024, 2048)
        FPM_3M_CASE(256, 1024, 4096)
        FPM_3M_CASE(512, 1024, 1024)
        FPM_3M_CASE(1024, 1024, 1024)
        FPM_3M_CASE(2048, 2048, 2048)
        FPM_3M_CASE(4096, 4096, 4096)
        default: throw HipError{hipErrorInvalidValue, "3 x 2^k transform: unsupported length"};
    }
#undef FPM_3M_CASE
    FPM_LAUNCH_CHECK();
}

const float2* generic_twiddles(uint32_t len) {
    static std::mutex mu;
    static std::map<std::pair<int, uint32_t>, float2*> tables;
    int dev = 0;
    FPM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    auto it = tables.find({dev, len});
    if (it != tables.end()) return it->second;
    std::vector<float2> h(len);
    const double two_pi = 6.283185307179586476925286766559;
    for (uint32_t k = 0; k < len; ++k) {
        const double ang = -two_pi * (double)k / (double)len;
        h[k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    float2* d = nullptr;
    FPM_HIP(hipMalloc(&d, sizeof(float2) * len));
    FPM_HIP(hipMemcpy(d, h.data(), sizeof(float2) * len, hipMemcpyHostToDevice));
    tables[{dev, len}] = d;
    return d;
}

void set_factors(GenFFT& a, uint32_t len) {
    uint32_t m = len;
    a.nf = 0;
    auto take = [&](uint32_t r) {
        while (m % r == 0) {
            if (a.nf >= (uint32_t)kGenMaxFactors) throw HipError{hipErrorInvalidValue, "generic FFT: too many factors"};
            a.fac[a.nf++] = (uint16_t)r;
            m /= r;
        }
    };
    take(4);
    take(2);
    take(3);
    // 5, 7 hard-coded; other primes summed directly
    for (uint32_t p = 5; p <= len; p += 2) {
        if (m == 1) break;
        take(p);
    }
}

// the plan of a length: factors, twiddles, divisors (not tl)
void plan_generic(GenFFT& a) {
    if (a.len < 2 || a.len > (uint32_t)kGenericMaxSide)
        throw HipError{hipErrorInvalidValue, "generic FFT: length out of range"};
    set_factors(a, a.len);
    a.tw = generic_twiddles(a.len);
    a.f_lpb = make_fastdiv(a.lines_per_batch ? a.lines_per_batch : 1);
    a.f_len = make_fastdiv(a.len);
    uint32_t s = 1;
    for (uint32_t f = 0; f < a.nf; ++f) {
        a.f_s[f] = make_fastdiv(s);
        a.f_nb[f] = make_fastdiv(a.len / a.fac[f]);
        a.f_r[f] = make_fastdiv(a.fac[f]);
        s *= a.fac[f];
    }
}

// threads of a workgroup holding pts points: about dens points each (at
// least 256 threads, at most kGenPts points each), a multiple of the wave
uint32_t gen_threads(uint32_t pts, uint32_t dens = 8) {
    const uint32_t threads =
        std::min<uint32_t>(FPM_GEN_LB, std::max<uint32_t>(256, (pts + dens * 64 - 1) / (dens * 64) * 64));
    if ((size_t)threads * kGenPts < pts) throw HipError{hipErrorInvalidValue, "generic FFT: line set too long"};
    return threads;
}

// ---- launch shapes of the transform passes and their autotuner
// A pass's shape is the points per workgroup (tl = max(1, pts / len) lines)
// and the points per thread.  Neither changes any result: every butterfly,
// twiddle and crop product is computed the same way whichever workgroup
// and thread hold it, so a shape is chosen for speed alone.  Defaults
// (measured on the 48-tile solve with 16 tiles of 192 px): contiguous
// lines 1024 points (4096 -> 1024: 39.6 -> 35.8 ms per epoch; 512: 37.2),
// strided lines 2048 (4096 -> 2048: 33.5 -> 32.2; 1024: 35.6), 8 points per
// thread.  generic_tune replaces them per device and (length, line kind,
// crop source) with the fastest of a candidate set timed on a solve's own
// buffers at its set-up.
struct GenShape {
    uint32_t pts, dens;
};
struct GenKey {
    int dev;
    uint32_t len;
    bool strided, crop;
    bool operator<(const GenKey& o) const {
        return std::tie(dev, len, strided, crop) < std::tie(o.dev, o.len, o.strided, o.crop);
    }
};
std::mutex g_shape_mu;
std::map<GenKey, GenShape> g_shapes;

GenKey key_of(const GenFFT& a, bool crop) {
    int dev = 0;
    FPM_HIP(hipGetDevice(&dev));
    return GenKey{dev, a.len, a.es != 1, crop};
}
GenShape shape_of(const GenFFT& a, bool crop) {
    if (FPM_GEN_TUNE) {
        std::lock_guard<std::mutex> lock(g_shape_mu);
        auto it = g_shapes.find(key_of(a, crop));
        if (it != g_shapes.end()) return it->second;
    }
    return GenShape{a.es == 1 ? (uint32_t)FPM_GEN_CONTIG_POINTS : (uint32_t)FPM_GEN_STRIDED_POINTS, 8};
}

void launch_generic_lines(GenFFT a, hipStream_t st, const GenCrop* crop = nullptr, const GenShape* force = nullptr) {
    plan_generic(a);
    if (const uint32_t m = len3m(a.len, a.es != 1)) {
        const GenCrop c = crop ? *crop : GenCrop{};
        if (a.inverse)
            crop ? launch_3m<true, true>(m, a, c, st) : launch_3m<true, false>(m, a, c, st);
        else
            crop ? launch_3m<false, true>(m, a, c, st) : launch_3m<false, false>(m, a, c, st);
        return;
    }
    const GenShape sh = force ? *force : shape_of(a, crop != nullptr);
    a.tl = std::max<uint32_t>(1, sh.pts / a.len);
    a.f_tl = make_fastdiv(a.tl);
    const uint32_t pts = a.tl * a.len;
    const uint32_t threads = gen_threads(pts, sh.dens);
    const size_t shmem = (size_t)pts * sizeof(float2);
    static std::once_flag big_lds;
    if (FPM_GEN_SET_ATTR) std::call_once(big_lds, [] {
        // lines up to 16384 points (128 KiB) stage in LDS
        FPM_HIP(hipFuncSetAttribute((const void*)fft_generic_kernel<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)(kGenericMaxSide * sizeof(float2))));
        FPM_HIP(hipFuncSetAttribute((const void*)fft_generic_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)(kGenericMaxSide * sizeof(float2))));
    });
    const uint32_t blocks = (a.nlines + a.tl - 1) / a.tl;
    if (crop)
        hipLaunchKernelGGL(fft_generic_kernel<true>, dim3(blocks), dim3(threads), shmem, st, a, *crop);
    else
        hipLaunchKernelGGL(fft_generic_kernel<false>, dim3(blocks), dim3(threads), shmem, st, a, GenCrop{});
    FPM_LAUNCH_CHECK();
}

// ------------------------------------------------------- solve kernels
// Every kernel runs one tile per grid row: tab (optional) gives grouped
// tiles, tile k = blockIdx.y with its pages, spectra, pupil and moments at
// block k of each array and its LED windows, images, mask and loss slot
// from tab[k]; without tab, grid row 0 and the arguments as given.

// crop x pupil into ifftshifted pages (kernels.cuh:38-64): page l pixel
// (di, dj) = S(window of LED led0 + l at pupil pixel (i, j)) P(i, j),
// i = (di + nr/2) mod nr, j = (dj + nc/2) mod nc; pages column-major, natural.
__global__ __launch_bounds__(256) void gen_crop_kernel(const float2* __restrict__ S, const float2* __restrict__ P,
                                                       const int2* __restrict__ led, uint32_t led0, uint32_t b,
                                                       uint32_t nr, uint32_t nc, uint32_t Nr, uint32_t Nc,
                                                       float2* __restrict__ out, const TileIn* __restrict__ tab) {
    const size_t plane = (size_t)nr * nc, total = plane * b;
    if (tab) {
        const uint32_t tk = blockIdx.y;
        S += (size_t)tk * Nr * Nc;
        P += (size_t)tk * plane;
        out += (size_t)tk * total;
        led = tab[tk].led;
    }
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < total; k += (size_t)gridDim.x * 256) {
        const uint32_t l = (uint32_t)(k / plane);
        const uint32_t pix = (uint32_t)(k - (size_t)l * plane);
        const uint32_t dj = pix / nr, di = pix - dj * nr;
        const uint32_t i = wrap1(di + nr / 2, nr), j = wrap1(dj + nc / 2, nc);
        const int2 kk = led[led0 + l];
        const uint32_t srow = wrap1(i + (uint32_t)kk.y - 1u + Nr / 2, Nr);
        const uint32_t scol = wrap1(j + (uint32_t)kk.x - 1u + Nc / 2, Nc);
        const float2 pp = P[(size_t)j * nr + i];
        float2 z = make_float2(0.f, 0.f);
        if (pp.x != 0.0f || pp.y != 0.0f) z = cmul_g(S[(size_t)scol * Nr + srow], pp);
        out[k] = z;
    }
}

// e = |z| / (n_r n_c r^2) - y of every pixel of the b pages
__global__ __launch_bounds__(256) void gen_residual_kernel(const float2* __restrict__ z, YStack y, uint32_t led0,
                                                           size_t plane, size_t total, float ratio, float inv_ratio,
                                                           float* __restrict__ e, const TileIn* __restrict__ tab) {
    if (tab) {
        const uint32_t tk = blockIdx.y;
        z += (size_t)tk * total;
        e += (size_t)tk * total;
        const TileIn ti = tab[tk];
        y.data = ti.y;
        y.mn = ti.ymn;
        y.range = ti.yrange;
        y.rcp_range = ti.yrcp;
    }
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < total; k += (size_t)gridDim.x * 256) {
        const float2 v = z[k];
        const float mag = sqrtf(v.x * v.x + v.y * v.y);
        e[k] = scaled_mag(mag, ratio, inv_ratio) - read_y(y, (size_t)led0 * plane + k);
    }
}

// the loss stencil (kernels.cuh:197-275) with cyclic neighbours:
// g = sign(e - e_left) - sign(e_right - e) + sign(e - e_up) - sign(e_down - e)
// (copysign semantics), u = r^2 g z/|z| (0 where |z| = 0), loss += |Dx e| + |Dy e|
__global__ __launch_bounds__(256) void gen_stencil_kernel(const float2* __restrict__ z, const float* __restrict__ e,
                                                          uint32_t nr, uint32_t nc, size_t total, float gscale,
                                                          float2* __restrict__ u, double* __restrict__ loss_acc,
                                                          const TileIn* __restrict__ tab, uint8_t* __restrict__ signs) {
    __shared__ double red[4];
    if (tab) {
        const uint32_t tk = blockIdx.y;
        z += (size_t)tk * total;
        e += (size_t)tk * total;
        u += (size_t)tk * total;
        loss_acc = tab[tk].loss;
    }
    const size_t plane = (size_t)nr * nc;
    float loss32 = 0.0f;
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < total; k += (size_t)gridDim.x * 256) {
        const size_t l = k / plane;
        const uint32_t pix = (uint32_t)(k - l * plane);
        const uint32_t j = pix / nr, i = pix - j * nr;
        const float* el = e + l * plane;
        const float e0 = el[(size_t)j * nr + i];
        const uint32_t jl = j == 0 ? nc - 1 : j - 1, jr = j + 1 == nc ? 0 : j + 1;
        const uint32_t iu = i == 0 ? nr - 1 : i - 1, id = i + 1 == nr ? 0 : i + 1;
        const float eL = el[(size_t)jl * nr + i];
        const float eR = el[(size_t)jr * nr + i];
        const float eU = el[(size_t)j * nr + iu];
        const float eD = el[(size_t)j * nr + id];
        const float gval =
            (copysignf(1.0f, e0 - eL) - copysignf(1.0f, eR - e0) + copysignf(1.0f, e0 - eU) - copysignf(1.0f, eD - e0)) *
            gscale;
        loss32 += fabsf(eR - e0) + fabsf(eD - e0);
        if (signs) signs[k] = (uint8_t)((signbit(eR - e0) ? 1 : 0) | (signbit(eD - e0) ? 2 : 0));
        const float2 v = z[k];
        const float mag = sqrtf(v.x * v.x + v.y * v.y);
        float2 o = make_float2(0.f, 0.f);
        if (mag > 0.0f) {
            const float inv = 1.0f / mag;
            o = make_float2(v.x * inv * gval, v.y * inv * gval);
        }
        u[k] = o;
    }
    if (loss_acc) {  // uniform over the workgroup
        double loss = (double)loss32;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = loss;
        __syncthreads();
        if (threadIdx.x == 0) unsafeAtomicAdd(loss_acc, red[0] + red[1] + red[2] + red[3]);
    }
}

// The per-LED column steps in one kernel: a workgroup holds lt whole
// columns (n_r points each) of one page in LDS -- columns j0 - 1 .. j0 + lt
// - 2 (cyclic), the first and last as neighbours only -- and runs the
// inverse column transform, the residual e (in LDS beside the columns), the
// loss stencil of gen_stencil_kernel on its lt - 2 interior columns, and the
// forward column transform, then writes the interior columns.  The row
// transforms (the crop's inverse pass before, the forward pass after) run
// as separate passes.  Grid: x = page (b per tile) x column group, y = tile.
struct GenCols {
    const float2* z;
    float2* out;
    YStack y;
    const TileIn* tab;
    double* loss;
    uint32_t led0, b, nr, nc, lt, groups;
    float ratio, inv_ratio, gscale;
};
__global__ __launch_bounds__(FPM_GEN_LB) void gen_cols_fused_kernel(GenFFT a, GenCols c) {
    extern __shared__ float2 sm[];
    __shared__ double red[FPM_GEN_LB / 64];
    const uint32_t lt = c.lt, len = c.nr, tot = lt * len;
    float2* x = sm;
    float* ef = (float*)(sm + tot);
    const uint32_t page = blockIdx.x / c.groups, grp = blockIdx.x - page * c.groups;
    const size_t plane = (size_t)c.nr * c.nc;
    size_t pbase = (size_t)page * plane;
    YStack y = c.y;
    double* loss_acc = c.loss;
    if (c.tab) {
        const uint32_t tk = blockIdx.y;
        pbase += (size_t)tk * c.b * plane;
        const TileIn ti = c.tab[tk];
        y.data = ti.y;
        y.mn = ti.ymn;
        y.range = ti.yrange;
        y.rcp_range = ti.yrcp;
        loss_acc = ti.loss;
    }
    const float2* __restrict__ z = c.z + pbase;
    float2* __restrict__ out = c.out + pbase;
    const size_t ybase = (size_t)(c.led0 + page) * plane;
    const uint32_t j0 = grp * (lt - 2);  // first interior column; line l is column j0 - 1 + l (mod n_c)
    auto column = [&](uint32_t line) -> uint32_t {
        const uint32_t v = j0 + line + c.nc - 1u;  // lt <= n_c + 2: v < 3 n_c
        return v >= c.nc ? (v - c.nc >= c.nc ? v - 2u * c.nc : v - c.nc) : v;
    };
    auto interior = [&](uint32_t line) -> bool { return line >= 1 && line + 1 < lt && j0 + line - 1u < c.nc; };
    for (uint32_t idx = threadIdx.x; idx < tot; idx += blockDim.x) {
        const uint32_t l = fdiv(idx, a.f_len), e = idx - l * len;
        x[idx] = z[(size_t)column(l) * len + e];
    }
    __syncthreads();
    gen_passes(a, x, lt, true);
    // the image loads of GY of a thread's points are issued before their
    // conversions, and the uniform choices are taken outside the loop (the
    // operations of read_y and scaled_mag)
    with_y_mode(y, c.inv_ratio, [&](auto u16, auto quot) {
        constexpr bool U16 = decltype(u16)::value, Q = decltype(quot)::value;
        constexpr uint32_t GY = 8;
        for (uint32_t i0 = threadIdx.x; i0 < tot; i0 += GY * blockDim.x) {
            uint32_t raw[GY];
#pragma unroll
            for (uint32_t u = 0; u < GY; ++u) {
                const uint32_t idx = i0 + u * blockDim.x;
                raw[u] = 0;
                if (idx < tot) {
                    const uint32_t l = fdiv(idx, a.f_len), e = idx - l * len;
                    raw[u] = load_y_raw<U16>(y, ybase + (size_t)column(l) * len + e);
                }
            }
#pragma unroll
            for (uint32_t u = 0; u < GY; ++u) {
                const uint32_t idx = i0 + u * blockDim.x;
                if (idx < tot) {
                    const float2 v = x[idx];
                    const float mag = sqrtf(v.x * v.x + v.y * v.y);
                    const float sm = Q ? mag / c.ratio : mag * c.inv_ratio;
                    ef[idx] = sm - y_of_raw<U16>(raw[u], y);
                }
            }
        }
    });
    __syncthreads();
    float loss32 = 0.0f;
    for (uint32_t idx = threadIdx.x; idx < tot; idx += blockDim.x) {
        const uint32_t l = fdiv(idx, a.f_len), i = idx - l * len;
        if (!interior(l)) continue;
        const float e0 = ef[idx];
        const float eL = ef[idx - len], eR = ef[idx + len];
        const float eU = ef[l * len + (i == 0 ? len - 1 : i - 1)];
        const float eD = ef[l * len + (i + 1 == len ? 0 : i + 1)];
        const float gval =
            (copysignf(1.0f, e0 - eL) - copysignf(1.0f, eR - e0) + copysignf(1.0f, e0 - eU) - copysignf(1.0f, eD - e0)) *
            c.gscale;
        loss32 += fabsf(eR - e0) + fabsf(eD - e0);
        const float2 v = x[idx];
        const float mag = sqrtf(v.x * v.x + v.y * v.y);
        float2 o = make_float2(0.f, 0.f);
        if (mag > 0.0f) {
            const float inv = 1.0f / mag;
            o = make_float2(v.x * inv * gval, v.y * inv * gval);
        }
        x[idx] = o;  // only this thread reads x[idx] in this loop
    }
    __syncthreads();
    gen_passes(a, x, lt, false);
    for (uint32_t idx = threadIdx.x; idx < tot; idx += blockDim.x) {
        const uint32_t l = fdiv(idx, a.f_len), e = idx - l * len;
        if (interior(l)) out[(size_t)column(l) * len + e] = x[idx];
    }
    if (loss_acc) {  // uniform over the workgroup
        double loss = (double)loss32;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = loss;
        __syncthreads();
        if (threadIdx.x == 0) {
            double sum = 0.0;
            for (uint32_t w = 0; w < blockDim.x / 64; ++w) sum += red[w];
            unsafeAtomicAdd(loss_acc, sum);
        }
    }
}

// owner-computes stitch: G_S(row, col) = sum over the batch's LEDs, in LED
// order, of R_l(fftshifted window pixel) conj(P) / (N_r N_c) (kernels.cuh:
// 107-116 with the atomics replaced by a gather); every element written.
// A workgroup owns a 64-row x 16-column block (thread: one row, 4 columns),
// first keeps the LEDs whose windows meet the block (LDS bit mask), then
// sums only those.
constexpr uint32_t kGsTR = 64, kGsTC = 16;
// cols[i] = c0 | c1 << 16: the smallest column interval of pupil row i that
// holds every P != 0 and A != 0 (0: none); P = 0 outside it at every step of
// a solve (the clamped update keeps P = 0 where A = 0).  Grouped tiles: grid
// row k is tile k (its pupil block, its A from tab[k], its nr words).
__global__ __launch_bounds__(256) void gen_pupil_cols_kernel(const float2* __restrict__ P, const float* __restrict__ A,
                                                             uint32_t nr, uint32_t nc, uint32_t* __restrict__ cols,
                                                             const TileIn* __restrict__ tab) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (tab) {
        const uint32_t tk = blockIdx.y;
        P += (size_t)tk * nr * nc;
        A = tab[tk].A;
        cols += (size_t)tk * nr;
    }
    if (i >= nr) return;
    uint32_t lo = 0xffffffffu, hi = 0;
    for (uint32_t j = 0; j < nc; ++j) {
        const float2 p = P[(size_t)j * nr + i];
        if (p.x != 0.0f || p.y != 0.0f || A[(size_t)j * nr + i] != 0.0f) {
            lo = min(lo, j);
            hi = j + 1;
        }
    }
    cols[i] = hi ? lo | hi << 16 : 0u;
}

__global__ __launch_bounds__(256) void gen_stitch_kernel(const float2* __restrict__ R, const float2* __restrict__ P,
                                                         const int2* __restrict__ led, uint32_t led0, uint32_t b,
                                                         uint32_t nr, uint32_t nc, uint32_t Nr, uint32_t Nc,
                                                         float invN2, float2* __restrict__ Gs,
                                                         const TileIn* __restrict__ tab,
                                                         const uint32_t* __restrict__ pcols) {
    __shared__ unsigned long long hit[4];
    const size_t plane = (size_t)nr * nc;
    if (tab) {
        const uint32_t tk = blockIdx.y;
        R += (size_t)tk * b * plane;
        P += (size_t)tk * plane;
        Gs += (size_t)tk * Nr * Nc;
        led = tab[tk].led;
        if (pcols) pcols += (size_t)tk * nr;
    }
    const uint32_t tiles_r = (Nr + kGsTR - 1) / kGsTR;
    const uint32_t ra = (blockIdx.x % tiles_r) * kGsTR, ca = (blockIdx.x / tiles_r) * kGsTC;
    const uint32_t t = threadIdx.x, r = t & 63u, cg = t >> 6;
    const uint32_t row = ra + r;
    float2 acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = make_float2(0.f, 0.f);
    for (uint32_t c0 = 0; c0 < b; c0 += 256) {
        if (t < 4) hit[t] = 0ull;
        __syncthreads();
        const uint32_t me = c0 + t;
        if (me < b) {
            const int2 kk = led[led0 + me];
            const uint32_t wr = wrap1((uint32_t)kk.y - 1u + Nr / 2, Nr);
            const uint32_t wc = wrap1((uint32_t)kk.x - 1u + Nc / 2, Nc);
            // window rows [wr, wr + nr) and block rows [ra, ra + 64) (mod N) meet
            // iff one start lies inside the other range (conservative at the
            // last, partial block)
            const bool rows = (ra >= wr ? ra - wr : ra + Nr - wr) < nr || (wr >= ra ? wr - ra : wr + Nr - ra) < kGsTR;
            const bool cols = (ca >= wc ? ca - wc : ca + Nc - wc) < nc || (wc >= ca ? wc - ca : wc + Nc - ca) < kGsTC;
            if (rows && cols) atomicOr(&hit[t >> 6], 1ull << (t & 63u));
        }
        __syncthreads();
        for (int w = 0; w < 4; ++w) {
            unsigned long long m = hit[w];
            while (m) {
                const int bit = __ffsll(m) - 1;
                m &= m - 1;
                const uint32_t l = c0 + (uint32_t)w * 64u + (uint32_t)bit;
                const int2 kk = led[led0 + l];
                const uint32_t wr = wrap1((uint32_t)kk.y - 1u + Nr / 2, Nr);
                const uint32_t wc = wrap1((uint32_t)kk.x - 1u + Nc / 2, Nc);
                if (row >= Nr) continue;
                const uint32_t li = row >= wr ? row - wr : row + Nr - wr;
                if (li >= nr) continue;
                const uint32_t di = wrap1(li + nr / 2, nr);
                // P = 0 outside pupil row li's support interval [c0, c0 + cw);
                // inside it R is loaded beside P
                uint32_t c0 = 0, cw = nc;
                if (pcols) {
                    const uint32_t cc = pcols[li];
                    c0 = cc & 0xffffu;
                    cw = (cc >> 16) - c0;
                }
                // the four columns' R and P loads issued before their products
                // (without intervals R is read wherever P is)
                float2 Rq[4], Pq[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    Rq[q] = Pq[q] = make_float2(0.f, 0.f);
                    const uint32_t col = ca + cg * 4u + (uint32_t)q;
                    if (col >= Nc) continue;
                    const uint32_t lj = col >= wc ? col - wc : col + Nc - wc;
                    if (lj - c0 < cw) {
                        const uint32_t dj = wrap1(lj + nc / 2, nc);
                        Rq[q] = R[(size_t)l * plane + (size_t)dj * nr + di];
                        Pq[q] = P[(size_t)lj * nr + li];
                    }
                }
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (Pq[q].x != 0.0f || Pq[q].y != 0.0f) {  // P = 0 contributes +-0
                        const float2 gw = cmulc_g(Rq[q], Pq[q]);
                        acc[q].x = fmaf(gw.x, invN2, acc[q].x);
                        acc[q].y = fmaf(gw.y, invN2, acc[q].y);
                    }
            }
        }
        __syncthreads();
    }
    if (row < Nr) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t col = ca + cg * 4u + (uint32_t)q;
            if (col < Nc) Gs[(size_t)col * Nr + row] = acc[q];
        }
    }
}

// pupil gradient sum_l R_l conj(S_l) in LED order (kernels.cuh:117-128), then
// the RMSprop + clamp of kernels.cuh:384-410 (mP, A given) or GP = gradient
__global__ __launch_bounds__(256) void gen_pupil_kernel(const float2* __restrict__ R, const float2* __restrict__ S,
                                                        const int2* __restrict__ led, uint32_t led0, uint32_t b,
                                                        uint32_t nr, uint32_t nc, uint32_t Nr, uint32_t Nc,
                                                        float2* __restrict__ P, float* __restrict__ mP,
                                                        const float* __restrict__ A, RmsArgs rms,
                                                        float2* __restrict__ GP, const TileIn* __restrict__ tab) {
    const uint32_t plane = nr * nc;
    const uint32_t pix = blockIdx.x * 256u + threadIdx.x;
    if (pix >= plane) return;
    if (tab) {
        const uint32_t tk = blockIdx.y;
        R += (size_t)tk * b * plane;
        S += (size_t)tk * Nr * Nc;
        P += (size_t)tk * plane;
        if (mP) mP += (size_t)tk * plane;
        if (GP) GP += (size_t)tk * plane;
        led = tab[tk].led;
        A = tab[tk].A;
    }
    const uint32_t j = pix / nr, i = pix - j * nr;
    const uint32_t di = wrap1(i + nr / 2, nr), dj = wrap1(j + nc / 2, nc);
    float2 gp = make_float2(0.f, 0.f);
    for (uint32_t l = 0; l < b; ++l) {
        const float2 Rv = R[(size_t)l * plane + (size_t)dj * nr + di];
        const int2 kk = led[led0 + l];
        const uint32_t srow = wrap1(i + (uint32_t)kk.y - 1u + Nr / 2, Nr);
        const uint32_t scol = wrap1(j + (uint32_t)kk.x - 1u + Nc / 2, Nc);
        const float2 gq = cmulc_g(Rv, S[(size_t)scol * Nr + srow]);
        gp.x += gq.x;
        gp.y += gq.y;
    }
    if (mP) {
        const float2 p = P[pix];
        float mu = gp.x * gp.x + gp.y * gp.y;
        mu = rms.beta * mP[pix] + (1.0f - rms.beta) * mu;
        const float sq = rms_lr(rms) / (sqrtf(mu) + rms.eps);
        float2 w = p;
        w.x -= gp.x * sq;
        w.y -= gp.y * sq;
        const float a = A[pix];
        const float ang = atan2f(w.y, w.x);
        float ab = sqrtf(w.x * w.x + w.y * w.y);
        ab = a * fmaxf(fminf(ab, a + rms.clamp), a - rms.clamp);
        w.x = cosf(ang) * ab;
        w.y = sinf(ang) * ab;
        P[pix] = w;
        mP[pix] = mu;
    } else {
        GP[pix] = gp;
    }
}

// RMSprop of O (kernels.cuh:364-382): m = beta m + (1-beta)|G|^2,
// O -= lr G / (sqrt(m) + eps)
__global__ __launch_bounds__(256) void gen_update_kernel(const float2* __restrict__ G, float2* __restrict__ O,
                                                         float* __restrict__ mO, size_t n, RmsArgs rms) {
    const float lr = rms_lr(rms);
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256) {
        const float2 gr = G[k];
        float mu = gr.x * gr.x + gr.y * gr.y;
        mu = rms.beta * mO[k] + (1.0f - rms.beta) * mu;
        const float sq = lr / (sqrtf(mu) + rms.eps);
        float2 w = O[k];
        w.x -= gr.x * sq;
        w.y -= gr.y * sq;
        O[k] = w;
        mO[k] = mu;
    }
}

// The object update on columns of N = 3 M points: the last (contiguous)
// pass of G_O = IDFT_N(G_S) on the 3 x 2^k transform, RMSprop of O on its
// outputs (gen_update_kernel's arithmetic, kernels.cuh:364-382) and, with
// FWD, the first (contiguous) pass of DFT_N of the new O, written back into
// G -- as the power-of-two path's fused update does: G_O never reaches HBM
// and the next step's spectrum starts here.  Contiguous lines of one
// workgroup are one contiguous run of nl N elements.
template <int M, int P, bool FWD>
__global__ __launch_bounds__(3 * P / 16) void gen_update3m_kernel(float2* __restrict__ G, float2* __restrict__ O,
                                                                float* __restrict__ mO, uint32_t nlines, RmsArgs rms,
                                                                const float2* __restrict__ twN, Twiddles tw) {
    constexpr int N = 3 * M, NL = P / M, T = 3 * P / 16, LD = N + 1;
    static_assert(NL >= 1 && NL * N == 16 * T, "16 points per thread");
    __shared__ float2 lds[Lines3m<M, NL>::WORDS];
    __shared__ float2 twl[2 * kTwLo];
    const int t = threadIdx.x;
    tw = bind_twiddles(tw, twl, t, T);
    const uint32_t L0 = blockIdx.x * (uint32_t)NL;
    const uint32_t nl = min((uint32_t)NL, nlines - L0);
    const uint32_t tot = nl * (uint32_t)N;
    const bool full = nl == (uint32_t)NL;
    float2* __restrict__ g = G + (size_t)L0 * N;
    float2* __restrict__ o = O + (size_t)L0 * N;
    float* __restrict__ m = mO + (size_t)L0 * N;
    if (full) {
        float2 buf[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) buf[k] = g[t + k * T];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const uint32_t idx = (uint32_t)(t + k * T), l = idx / (uint32_t)N;
            lds[l * LD + (idx - l * (uint32_t)N)] = buf[k];
        }
    } else {
        for (uint32_t idx = t; idx < tot; idx += T) {
            const uint32_t l = idx / (uint32_t)N;
            lds[l * LD + (idx - l * (uint32_t)N)] = g[idx];
        }
    }
    __syncthreads();
    fft3m_lines<M, NL, true>(lds, nl, twN, tw, t);
    const float lr = rms_lr(rms);
    auto update = [&](uint32_t idx) {
        const uint32_t l = idx / (uint32_t)N, a = l * LD + (idx - l * (uint32_t)N);
        const float2 gr = lds[a];
        float mu = gr.x * gr.x + gr.y * gr.y;
        mu = rms.beta * m[idx] + (1.0f - rms.beta) * mu;
        const float sq = lr / (sqrtf(mu) + rms.eps);
        float2 w = o[idx];
        w.x -= gr.x * sq;
        w.y -= gr.y * sq;
        o[idx] = w;
        m[idx] = mu;
        lds[a] = w;  // only this thread touches lds[a] in this phase
    };
    if (full) {
#pragma unroll
        for (int k = 0; k < 16; ++k) update((uint32_t)(t + k * T));
    } else {
        for (uint32_t idx = t; idx < tot; idx += T) update(idx);
    }
    if constexpr (FWD) {
        __syncthreads();
        fft3m_lines<M, NL, false>(lds, nl, twN, tw, t);
        if (full) {
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const uint32_t idx = (uint32_t)(t + k * T), l = idx / (uint32_t)N;
                g[idx] = lds[l * LD + (idx - l * (uint32_t)N)];
            }
        } else {
            for (uint32_t idx = t; idx < tot; idx += T) {
                const uint32_t l = idx / (uint32_t)N;
                g[idx] = lds[l * LD + (idx - l * (uint32_t)N)];
            }
        }
    }
}

unsigned gblocks(size_t n) { return (unsigned)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 1u << 16)); }

}  // namespace

void generic_prepare(const Geom& g) {
    for (uint32_t len : {g.nr, g.nc, g.Nr, g.Nc}) (void)generic_twiddles(len);
}

namespace {
// the fastest shape of pass a (its output redirected to scratch), timed over
// three launches per candidate after one untimed launch
GenShape tune_pass(GenFFT a, const GenCrop* crop, float2* scratch, hipStream_t st) {
    if (len3m(a.len, a.es != 1)) return shape_of(a, crop != nullptr);  // fixed shape: nothing to tune
    a.out = scratch;
    std::vector<GenShape> cand;
    std::vector<std::pair<uint32_t, uint32_t>> seen;  // (lines per workgroup, threads)
    for (uint32_t pts : {512u, 1024u, 2048u, 4096u, 8192u, 16384u})
        for (uint32_t dens : {4u, 8u, 16u}) {
            const uint32_t tl = std::max<uint32_t>(1, pts / a.len);
            if ((size_t)tl * a.len > (size_t)kGenericMaxSide) continue;
            if (tl > 1 && tl > a.nlines) continue;
            const uint32_t p = tl * a.len;
            const uint32_t th = std::min<uint32_t>(FPM_GEN_LB, std::max<uint32_t>(256, (p + dens * 64 - 1) / (dens * 64) * 64));
            if ((size_t)th * kGenPts < p) continue;
            if (std::find(seen.begin(), seen.end(), std::make_pair(tl, th)) != seen.end()) continue;
            seen.emplace_back(tl, th);
            cand.push_back(GenShape{pts, dens});
        }
    hipEvent_t e0, e1;
    FPM_HIP(hipEventCreate(&e0));
    FPM_HIP(hipEventCreate(&e1));
    GenShape best = shape_of(a, crop != nullptr);
    float best_ms = 1e30f;
    for (const GenShape& c : cand) {
        launch_generic_lines(a, st, crop, &c);
        FPM_HIP(hipEventRecord(e0, st));
        for (int r = 0; r < 3; ++r) launch_generic_lines(a, st, crop, &c);
        FPM_HIP(hipEventRecord(e1, st));
        FPM_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        FPM_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best_ms) {
            best_ms = ms;
            best = c;
        }
    }
    FPM_HIP(hipEventDestroy(e0));
    FPM_HIP(hipEventDestroy(e1));
    return best;
}
}  // namespace

void generic_tune(const Geom& g, uint32_t T, uint32_t B, const float2* O, const float2* S, const float2* P,
                  const int2* led, const TileIn* tab, const float2* pages, hipStream_t st) {
    if (!FPM_GEN_TUNE) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    FPM_HIP(hipStreamIsCapturing(st, &cs));
    if (cs != hipStreamCaptureStatusNone) return;  // timing needs a live stream
    const bool fused = generic_fused_pages(g, false);
    const size_t NN = (size_t)g.Nr * g.Nc, nn = (size_t)g.nr * g.nc;
    // the passes of a step as the solve launches them (in: its own buffers)
    std::vector<std::pair<GenFFT, bool>> passes;  // (pass, crop source)
    GenFFT a{};
    a.in = O;  // DFT_N of the objects: contiguous lines, then strided
    a.len = g.Nr;
    a.nlines = g.Nc * T;
    a.lines_per_batch = a.nlines;
    a.line_stride = g.Nr;
    a.es = 1;
    passes.emplace_back(a, false);
    a = GenFFT{};
    a.in = S;
    a.len = g.Nc;
    a.nlines = g.Nr * T;
    a.lines_per_batch = g.Nr;
    a.line_stride = 1;
    a.es = g.Nr;
    a.batch_stride = NN;
    passes.emplace_back(a, false);
    a = GenFFT{};  // per-LED row passes (crop source, then the forward pass)
    a.in = pages;
    a.len = g.nc;
    a.nlines = g.nr * T * B;
    a.lines_per_batch = g.nr;
    a.line_stride = 1;
    a.es = g.nr;
    a.batch_stride = nn;
    a.inverse = 1;
    if (fused) passes.emplace_back(a, true);
    a.inverse = 0;
    passes.emplace_back(a, false);
    if (!fused) {  // per-LED column passes
        a = GenFFT{};
        a.in = pages;
        a.len = g.nr;
        a.nlines = g.nc * T * B;
        a.lines_per_batch = a.nlines;
        a.line_stride = g.nr;
        a.es = 1;
        passes.emplace_back(a, false);
    }
    GenCrop c{S, P, led, tab, 0, B, g.nr, g.nc, g.Nr, g.Nc, make_fastdiv(B)};
    std::vector<std::pair<GenKey, std::pair<GenFFT, bool>>> todo;
    {
        std::lock_guard<std::mutex> lock(g_shape_mu);
        for (auto& pc : passes) {
            const GenKey k = key_of(pc.first, pc.second);
            if (g_shapes.count(k) == 0 &&
                std::none_of(todo.begin(), todo.end(), [&](const auto& t) { return !(t.first < k) && !(k < t.first); }))
                todo.emplace_back(k, pc);
        }
    }
    if (todo.empty()) return;
    size_t most = 0;
    for (auto& t : todo) most = std::max<size_t>(most, (size_t)t.second.first.nlines * t.second.first.len);
    float2* scratch = nullptr;
    FPM_HIP(hipMalloc(&scratch, most * sizeof(float2)));
    try {
        for (auto& t : todo) {
            const GenShape best = tune_pass(t.second.first, t.second.second ? &c : nullptr, scratch, st);
            std::lock_guard<std::mutex> lock(g_shape_mu);
            g_shapes[t.first] = best;
        }
    } catch (...) {
        (void)hipFree(scratch);
        throw;
    }
    FPM_HIP(hipStreamSynchronize(st));
    FPM_HIP(hipFree(scratch));
}

void generic_strided_pass(float2* data, uint32_t rows, uint32_t cols, uint32_t batch, bool inverse, hipStream_t st) {
    GenFFT c{};
    c.in = data;
    c.out = data;
    c.len = cols;
    c.nlines = rows * batch;
    c.lines_per_batch = rows;
    c.line_stride = 1;
    c.es = rows;
    c.batch_stride = (uint64_t)rows * cols;
    c.inverse = inverse ? 1 : 0;
    launch_generic_lines(c, st);
}

bool generic_update_fusable(const Geom& g) { return FPM_GEN_FUSED_UPDATE && g.generic && len3m(g.Nr, false) != 0; }

void generic_update_fused(float2* G, float2* O, float* mO, const Geom& g, uint32_t ntiles, RmsArgs rms, bool forward,
                          hipStream_t st) {
    const uint32_t m = len3m(g.Nr, false);
    if (!m) throw HipError{hipErrorInvalidValue, "fused generic update: N_rows is not 3 x 2^k"};
    const float2* twN = generic_twiddles(g.Nr);
    int dev = 0;
    FPM_HIP(hipGetDevice(&dev));
    const Twiddles tw{twiddle_table(dev)};
    const uint32_t nlines = g.Nc * ntiles;
    auto go = [&](auto mm, auto pp) {
        constexpr int MM = decltype(mm)::value, PP = decltype(pp)::value, NL = PP / MM;
        const dim3 grid((nlines + NL - 1) / NL);
        if (forward)
            hipLaunchKernelGGL((gen_update3m_kernel<MM, PP, true>), grid, dim3(3 * PP / 16), 0, st, G, O, mO, nlines,
                               rms, twN, tw);
        else
            hipLaunchKernelGGL((gen_update3m_kernel<MM, PP, false>), grid, dim3(3 * PP / 16), 0, st, G, O, mO, nlines,
                               rms, twN, tw);
    };
    using std::integral_constant;
    switch (m) {
        case 16: go(integral_constant<int, 16>{}, integral_constant<int, 1024>{}); break;
        case 32: go(integral_constant<int, 32>{}, integral_constant<int, 1024>{}); break;
        case 64: go(integral_constant<int, 64>{}, integral_constant<int, 1024>{}); break;
        case 128: go(integral_constant<int, 128>{}, integral_constant<int, 1024>{}); break;
        case 256: go(integral_constant<int, 256>{}, integral_constant<int, 1024>{}); break;
        case 512: go(integral_constant<int, 512>{}, integral_constant<int, 1024>{}); break;
        case 1024: go(integral_constant<int, 1024>{}, integral_constant<int, 1024>{}); break;
        case 2048: go(integral_constant<int, 2048>{}, integral_constant<int, 2048>{}); break;
        case 4096: go(integral_constant<int, 4096>{}, integral_constant<int, 4096>{}); break;
        default: throw HipError{hipErrorInvalidValue, "fused generic update: unsupported length"};
    }
    FPM_LAUNCH_CHECK();
}

void fft2_generic(const float2* in, float2* out, uint32_t rows, uint32_t cols, uint32_t batch, bool inverse,
                  hipStream_t st) {
    // along the contiguous index: one line per column of every page
    GenFFT a{};
    a.in = in;
    a.out = out;
    a.len = rows;
    a.nlines = cols * batch;
    a.lines_per_batch = a.nlines;
    a.line_stride = rows;
    a.es = 1;
    a.batch_stride = 0;
    a.inverse = inverse ? 1 : 0;
    launch_generic_lines(a, st);
    // along the strided index: one line per row of every page, in place
    GenFFT c{};
    c.in = out;
    c.out = out;
    c.len = cols;
    c.nlines = rows * batch;
    c.lines_per_batch = rows;
    c.line_stride = 1;
    c.es = rows;
    c.batch_stride = (uint64_t)rows * cols;
    c.inverse = a.inverse;
    launch_generic_lines(c, st);
}

void generic_crop(const float2* S, const float2* P, const int2* led, uint32_t led0, uint32_t b, const Geom& g,
                  float2* out, hipStream_t st, const TileIn* tab, uint32_t ntiles) {
    if (!tab) ntiles = 1;
    hipLaunchKernelGGL(gen_crop_kernel, dim3(gblocks((size_t)g.nr * g.nc * b), ntiles), dim3(256), 0, st, S, P, led,
                       led0, b, g.nr, g.nc, g.Nr, g.Nc, out, tab);
    FPM_LAUNCH_CHECK();
}

void generic_stencil(const float2* z, const YStack& y, uint32_t led0, uint32_t b, int pratio, const Geom& g,
                     float* ebuf, float2* u, double* loss_acc, hipStream_t st, const TileIn* tab, uint32_t ntiles,
                     uint8_t* signs) {
    if (tab && signs) throw HipError{hipErrorInvalidValue, "generic stencil: sign output is not available for tiles"};
    if (!tab) ntiles = 1;
    const size_t plane = (size_t)g.nr * g.nc, total = plane * b;
    const float ratio = (float)(g.plane() * (uint32_t)(pratio * pratio));
    const float inv_ratio = exact_inverse_or_zero(ratio);
    hipLaunchKernelGGL(gen_residual_kernel, dim3(gblocks(total), ntiles), dim3(256), 0, st, z, y, led0, plane, total,
                       ratio, inv_ratio, ebuf, tab);
    FPM_LAUNCH_CHECK();
    hipLaunchKernelGGL(gen_stencil_kernel, dim3(gblocks(total), ntiles), dim3(256), 0, st, z, ebuf, g.nr, g.nc, total,
                       (float)(pratio * pratio), u, loss_acc, tab, signs);
    FPM_LAUNCH_CHECK();
}

// columns per workgroup of the fused column steps: about 4096 points (at
// least 8 columns: a quarter or less halo) and at most 12288 points (12 B
// each in LDS: the columns and their residuals); 0 when 3 columns do not fit
uint32_t gen_cols_lines(const Geom& g) {
    uint32_t lt = std::max<uint32_t>(4096 / g.nr, std::min<uint32_t>(8, 12288 / g.nr));
    lt = std::min<uint32_t>(lt, g.nc + 2);
    return lt >= 3 ? lt : 0;
}

bool generic_fused_pages(const Geom& g, bool signs) {
    return FPM_GEN_FUSED && !signs && g.nr <= (uint32_t)FPM_GEN_FUSED_MAX && gen_cols_lines(g) != 0;
}

void generic_inverse_pages(const float2* S, const float2* P, const int2* led, uint32_t led0, uint32_t b,
                           const Geom& g, float2* lat, hipStream_t st, const TileIn* tab, uint32_t ntiles,
                           bool fused) {
    if (!tab) ntiles = 1;
    if (!fused) {
        generic_crop(S, P, led, led0, b, g, lat, st, tab, ntiles);
        fft2_generic(lat, lat, g.nr, g.nc, ntiles * b, true, st);
        return;
    }
    // the row pass with the crop as its source, natural pages out
    GenFFT r{};
    r.out = lat;
    r.len = g.nc;
    r.nlines = g.nr * ntiles * b;
    r.lines_per_batch = g.nr;
    r.line_stride = 1;
    r.es = g.nr;
    r.batch_stride = (uint64_t)g.nr * g.nc;
    r.inverse = 1;
    GenCrop c{S, P, led, tab, led0, b, g.nr, g.nc, g.Nr, g.Nc, make_fastdiv(b)};
    launch_generic_lines(r, st, &c);
}

void generic_stencil_pages(const float2* z, const YStack& y, uint32_t led0, uint32_t b, int pratio, const Geom& g,
                           float* ebuf, float2* u, double* loss_acc, hipStream_t st, const TileIn* tab,
                           uint32_t ntiles, uint8_t* signs, bool fused) {
    if (!fused) {
        generic_stencil(z, y, led0, b, pratio, g, ebuf, u, loss_acc, st, tab, ntiles, signs);
        return;
    }
    if (signs) throw HipError{hipErrorInvalidValue, "generic fused columns: no sign output"};
    if (!tab) ntiles = 1;
    GenFFT a{};
    a.len = g.nr;
    a.lines_per_batch = 1;
    plan_generic(a);
    GenCols c{};
    c.z = z;
    c.out = u;
    c.y = y;
    c.tab = tab;
    c.loss = loss_acc;
    c.led0 = led0;
    c.b = b;
    c.nr = g.nr;
    c.nc = g.nc;
    c.lt = gen_cols_lines(g);
    c.groups = (g.nc + c.lt - 3) / (c.lt - 2);
    c.ratio = (float)(g.plane() * (uint32_t)(pratio * pratio));
    c.inv_ratio = exact_inverse_or_zero(c.ratio);
    c.gscale = (float)(pratio * pratio);
    const uint32_t pts = c.lt * g.nr;
    const size_t shmem = (size_t)pts * (sizeof(float2) + sizeof(float));
    static std::once_flag big_lds;
    std::call_once(big_lds, [] {
        FPM_HIP(hipFuncSetAttribute((const void*)gen_cols_fused_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)(12288 * (sizeof(float2) + sizeof(float)))));
    });
    if ((size_t)b * c.groups > 0x7fffffffu) throw HipError{hipErrorInvalidValue, "generic fused columns: grid too large"};
    hipLaunchKernelGGL(gen_cols_fused_kernel, dim3(b * c.groups, ntiles), dim3(gen_threads(pts)), shmem, st, a, c);
    FPM_LAUNCH_CHECK();
}

void generic_forward_pages(float2* rec, uint32_t pages, const Geom& g, hipStream_t st, bool fused) {
    if (!fused) {
        fft2_generic(rec, rec, g.nr, g.nc, pages, false, st);
        return;
    }
    GenFFT r{};
    r.in = rec;
    r.out = rec;
    r.len = g.nc;
    r.nlines = g.nr * pages;
    r.lines_per_batch = g.nr;
    r.line_stride = 1;
    r.es = g.nr;
    r.batch_stride = (uint64_t)g.nr * g.nc;
    r.inverse = 0;
    launch_generic_lines(r, st);
}

void generic_pupil_cols(const float2* P, const float* A, const Geom& g, uint32_t* cols, hipStream_t st,
                        const TileIn* tab, uint32_t ntiles) {
    if (!tab) ntiles = 1;
    if (g.nc > 65535u) throw HipError{hipErrorInvalidValue, "pupil column intervals: pupil side above 65535"};
    hipLaunchKernelGGL(gen_pupil_cols_kernel, dim3((g.nr + 255) / 256, ntiles), dim3(256), 0, st, P, A, g.nr, g.nc,
                       cols, tab);
    FPM_LAUNCH_CHECK();
}

void generic_stitch(const float2* R, const float2* P, const int2* led, uint32_t led0, uint32_t b, const Geom& g,
                    float2* Gs, hipStream_t st, const TileIn* tab, uint32_t ntiles, const uint32_t* pcols) {
    if (!tab) ntiles = 1;
    const uint32_t blocks = ((g.Nr + kGsTR - 1) / kGsTR) * ((g.Nc + kGsTC - 1) / kGsTC);
    hipLaunchKernelGGL(gen_stitch_kernel, dim3(blocks, ntiles), dim3(256), 0, st, R, P, led, led0, b, g.nr, g.nc,
                       g.Nr, g.Nc, 1.0f / (float)(g.Nr * g.Nc), Gs, tab, pcols);
    FPM_LAUNCH_CHECK();
}

void generic_pupil(const float2* R, const float2* S, const int2* led, uint32_t led0, uint32_t b, const Geom& g,
                   float2* P, float* mP, const float* A, RmsArgs rms, float2* GP, hipStream_t st, const TileIn* tab,
                   uint32_t ntiles) {
    if (!tab) ntiles = 1;
    hipLaunchKernelGGL(gen_pupil_kernel, dim3((g.plane() + 255) / 256, ntiles), dim3(256), 0, st, R, S, led, led0, b,
                       g.nr, g.nc, g.Nr, g.Nc, P, mP, A, rms, GP, tab);
    FPM_LAUNCH_CHECK();
}

void generic_update(const float2* G, float2* O, float* mO, size_t n, RmsArgs rms, hipStream_t st) {
    hipLaunchKernelGGL(gen_update_kernel, dim3(gblocks(n)), dim3(256), 0, st, G, O, mO, n, rms);
    FPM_LAUNCH_CHECK();
}

}  // namespace fpm
