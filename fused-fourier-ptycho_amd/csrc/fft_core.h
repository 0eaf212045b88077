// fft_core.h -- device building blocks of the HIP FFTs (gfx950).
//
// Power-of-two Stockham (self-sorting) transforms.  A line of LEN points is
// split over LEN/16 "butterfly slots"; each thread keeps 16 points in
// registers, runs a radix-16 (or one smaller radix) butterfly per pass and
// exchanges through LDS between passes, so the first pass reads global
// memory and the last pass writes global memory directly, both coalesced.
//
// Pass with span L (points already combined) and radix R, butterfly j:
//   in : x[j + s*LEN/R]                       s = 0..R-1
//   tw : x_s *= W_{L R}^{s (j mod L)}         (W_M = exp(-2 pi i / M), conj for inverse)
//   out: y[(j / L) L R + (j mod L) + s L]
// The radix list is 16,16,...,16,r for forward transforms and r,16,...,16 for
// inverse ones (r in {2,4,8} when log2 LEN is not a multiple of 4), so the
// last inverse pass and the first forward pass use the same slot mapping:
// a kernel can finish an inverse transform and start a forward one on the
// same registers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"

namespace fpm {

// ------------------------------------------------------------------ complex
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// Complex products with their roundings written out (each part: the first
// product fused onto the rounded second), so that every kernel -- and every
// slot map of one transform -- gets the same bits whatever the compiler
// would contract in its context.
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.y, b.x, -(a.x * b.y)));
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// a * (-i) for the forward direction, a * (+i) for the inverse one
template <bool INV> __device__ __forceinline__ float2 mul_mi(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// largest transform length handled by one kernel pass / the twiddle table
constexpr int kTwiddleLen = 16384;
constexpr int kTwiddleLog = 14;

// W_{kTwiddleLen}^m.  `tab` is the FP64-generated, FP32-rounded table in
// global memory (kernel argument); inside a kernel the factors are read from a
// two-level copy in LDS, W^m = W^(128 h) W^l with m = 128 h + l (256 entries,
// 2 KiB), so a twiddle costs two LDS reads and one complex product (<= 2 ulp
// beyond the table's rounding) instead of an L2 gather from a 128 KiB table.
constexpr int kTwLo = 128;
struct Twiddles {
    const float2* __restrict__ tab;   // global table, kTwiddleLen entries
    const float2* lds = nullptr;      // [0,128): W^(128 h), [128,256): W^l
    template <bool INV> __device__ __forceinline__ float2 get(uint32_t m) const {
        m &= (kTwiddleLen - 1);
        const float2 w = cmul(lds[m >> 7], lds[kTwLo + (m & (kTwLo - 1))]);
        return INV ? cconj(w) : w;
    }
};

// Fills the LDS copy (256 words) and binds it; ends with a barrier.
__device__ __forceinline__ Twiddles bind_twiddles(Twiddles tw, float2* lds_tw, int t, int T) {
    for (int k = t; k < 2 * kTwLo; k += T)
        lds_tw[k] = k < kTwLo ? tw.tab[k * (kTwiddleLen / kTwLo)] : tw.tab[k - kTwLo];
    __syncthreads();
    tw.lds = lds_tw;
    return tw;
}

// W^(s m) for s = 1..R-1 from 4 lookups (m, 2m, 4m, 8m) and at most two
// products more (error <= 3 products of table-grade values)
template <int R, bool INV>
__device__ __forceinline__ void twiddle_powers(const Twiddles& tw, uint32_t m, float2 (&w)[16]) {
    w[1] = tw.get<INV>(m);
    if constexpr (R > 2) {
        w[2] = tw.get<INV>(2 * m);
        w[3] = cmul(w[1], w[2]);
    }
    if constexpr (R > 4) {
        w[4] = tw.get<INV>(4 * m);
        w[5] = cmul(w[4], w[1]);
        w[6] = cmul(w[4], w[2]);
        w[7] = cmul(w[4], w[3]);
    }
    if constexpr (R > 8) {
        w[8] = tw.get<INV>(8 * m);
#pragma unroll
        for (int s = 9; s < 16; ++s) w[s] = cmul(w[8], w[s - 8]);
    }
}

// ----------------------------------------------------------- small DFTs
template <bool INV> __device__ __forceinline__ void dft2(float2& a0, float2& a1) {
    float2 t = a0;
    a0 = cadd(t, a1);
    a1 = csub(t, a1);
}

template <bool INV> __device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
    float2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
    float2 t2 = cadd(a1, a3), t3 = mul_mi<INV>(csub(a1, a3));
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}

// radix-8 as 4 x 2: n = 2 n1 + n2, k = k1 + 4 k2.  The W8^(n2 k1) twiddles
// of k1 = 1 and 3 are h (1 -+ i) multiples; each is fused into its radix-2
// butterfly as fmaf(+-h, s, a) with the roundings written out (left to the
// compiler, whether the product was fused into the butterfly's add followed
// the surrounding code, so two builds of one transform could differ).
template <bool INV> __device__ __forceinline__ void dft8(float2* a) {
    constexpr float h = 0.70710678118654752440f;
    // DFT4 over n1 for each n2
    dft4<INV>(a[0], a[2], a[4], a[6]);
    dft4<INV>(a[1], a[3], a[5], a[7]);
    // k1 = 1: a3 <- a3 W8 = h (s1, s2); (a2, a3) <- (a2 + a3, a2 - a3)
    {
        const float x = a[3].x, y = a[3].y;
        const float s1 = INV ? x - y : x + y, s2 = INV ? x + y : y - x;
        const float2 p = a[2];
        a[2] = make_float2(fmaf(h, s1, p.x), fmaf(h, s2, p.y));
        a[3] = make_float2(fmaf(-h, s1, p.x), fmaf(-h, s2, p.y));
    }
    // k1 = 2: W8^2 = -+i, exact
    a[5] = mul_mi<INV>(a[5]);
    dft2<INV>(a[4], a[5]);
    // k1 = 3: a7 <- a7 W8^3 = h (s1, s2) with signs folded into +-h
    {
        const float x = a[7].x, y = a[7].y;
        const float2 p = a[6];
        if constexpr (INV) {  // W8^-3 = (-1 + i)/sqrt2: (-h (x + y), h (x - y))
            a[6] = make_float2(fmaf(-h, x + y, p.x), fmaf(h, x - y, p.y));
            a[7] = make_float2(fmaf(h, x + y, p.x), fmaf(-h, x - y, p.y));
        } else {  // W8^3 = (-1 - i)/sqrt2: (h (y - x), -h (x + y))
            a[6] = make_float2(fmaf(h, y - x, p.x), fmaf(-h, x + y, p.y));
            a[7] = make_float2(fmaf(-h, y - x, p.x), fmaf(h, x + y, p.y));
        }
    }
    dft2<INV>(a[0], a[1]);
    // result X[k1 + 4 k2] sits at a[2 k1 + k2]; reorder to natural
    float2 t1 = a[1], t2 = a[2], t3 = a[3], t4 = a[4], t5 = a[5], t6 = a[6];
    a[1] = t2; a[2] = t4; a[3] = t6; a[4] = t1; a[5] = t3; a[6] = t5;
}

// W16^m applied to a (forward sign; conj for the inverse)
template <bool INV> __device__ __forceinline__ float2 w16(float2 a, int m) {
    constexpr float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;
    constexpr float h = 0.70710678118654752440f;
    float2 w;
    switch (m) {
        case 1: w = make_float2(c1, -s1); break;
        case 2: w = make_float2(h, -h); break;
        case 3: w = make_float2(s1, -c1); break;
        case 4: w = make_float2(0.f, -1.f); break;
        case 6: w = make_float2(-h, -h); break;
        case 9: w = make_float2(-c1, s1); break;
        default: w = make_float2(1.f, 0.f); break;
    }
    if (INV) w.y = -w.y;
    return cmul(a, w);
}

// radix-16 as 4 x 4: n = 4 n1 + n2, k = k1 + 4 k2
template <bool INV> __device__ __forceinline__ void dft16(float2* a) {
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) dft4<INV>(a[n2], a[4 + n2], a[8 + n2], a[12 + n2]);
    // element (n2, k1) now at a[4 k1 + n2]; twiddle W16^(n2 k1)
    a[5] = w16<INV>(a[5], 1);
    a[6] = w16<INV>(a[6], 2);
    a[7] = w16<INV>(a[7], 3);
    a[9] = w16<INV>(a[9], 2);
    a[10] = mul_mi<INV>(a[10]);
    a[11] = w16<INV>(a[11], 6);
    a[13] = w16<INV>(a[13], 3);
    a[14] = w16<INV>(a[14], 6);
    a[15] = w16<INV>(a[15], 9);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) dft4<INV>(a[4 * k1], a[4 * k1 + 1], a[4 * k1 + 2], a[4 * k1 + 3]);
    // X[k1 + 4 k2] at a[4 k1 + k2]: transpose the 4x4
    float2 t;
    t = a[1]; a[1] = a[4]; a[4] = t;
    t = a[2]; a[2] = a[8]; a[8] = t;
    t = a[3]; a[3] = a[12]; a[12] = t;
    t = a[6]; a[6] = a[9]; a[9] = t;
    t = a[7]; a[7] = a[13]; a[13] = t;
    t = a[11]; a[11] = a[14]; a[14] = t;
}

template <int R, bool INV> __device__ __forceinline__ void dft(float2* a) {
    if constexpr (R == 2) dft2<INV>(a[0], a[1]);
    else if constexpr (R == 4) dft4<INV>(a[0], a[1], a[2], a[3]);
    else if constexpr (R == 8) dft8<INV>(a);
    else dft16<INV>(a);
}

// ------------------------------------------------------------ pass plans
constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v >> 1); }

template <int LEN, bool INV> struct Plan {
    static constexpr int LOG = ilog2(LEN);
    static constexpr int REM = LOG % 4;                 // bits of the small pass
    static constexpr int SMALL = REM ? (1 << REM) : 0;  // 0: none
    static constexpr int NPASS = LOG / 4 + (REM ? 1 : 0);
    // radix of pass p
    static constexpr int radix(int p) {
        if (!REM) return 16;
        if (INV) return p == 0 ? SMALL : 16;
        return p == NPASS - 1 ? SMALL : 16;
    }
    // span L before pass p
    static constexpr int span(int p) {
        int l = 1;
        for (int q = 0; q < p; ++q) l *= radix(q);
        return l;
    }
};

// Slot maps.  A kernel runs NLINES lines of LEN points with T threads and 16
// points per thread.  In a pass of radix R, thread t owns butterflies
// b = t + T*q (q < 16/R); Map::split(b, NB) names the line and butterfly j
// (NB = LEN/R butterflies per line) and Map::lds(line, idx) the LDS word.
//   LineSlow: consecutive threads walk along one line (contiguous lines).
//   LineFast: consecutive threads walk across lines (strided lines: adjacent
//             lines are adjacent in memory, 16 of them make 128 bytes).
template <int NLINES, int LEN> struct LineSlowMap {
    static constexpr int PADLEN = LEN + LEN / 16;  // one pad word per 16
    static constexpr int LDS_WORDS = NLINES * PADLEN;
    __device__ static __forceinline__ void split(int b, int NB, int& line, int& j) { line = b / NB; j = b % NB; }
    __device__ static __forceinline__ int lds(int line, int idx) { return line * PADLEN + idx + (idx >> 4); }
};
// With 8 lines a 16-lane group of an 8-byte LDS access touches two point
// indices -- idx and idx + 1 in the reads, idx and idx + 16 in a first pass's
// writes -- so the 8-line map has no pad and swaps the low bit of idx when bit
// 4 is set: the two indices then fill disjoint halves of the 16 bank pairs.
// Wider maps (16+ lines) keep one pad word per point index.
template <int NLINES, int LEN> struct LineFastMap {
    static constexpr bool SWZ = NLINES == 8;
    static constexpr int LDS_WORDS = SWZ ? LEN * NLINES : LEN * (NLINES + 1);
    __device__ static __forceinline__ void split(int b, int NB, int& line, int& j) { line = b % NLINES; j = b / NLINES; }
    __device__ static __forceinline__ int lds(int line, int idx) {
        if constexpr (SWZ) return ((idx & ~1) | ((idx ^ (idx >> 4)) & 1)) * NLINES + line;
        else return idx * (NLINES + 1) + line;
    }
};

// ------------------------------------------------------------------ engines
// ENG_F32:     every butterfly in FP32 on the vector ALUs (thread = butterfly).
// ENG_SPLIT16: the radix-16 DFTs run on the matrix cores as FP16 products
//              with split hi + lo operands and FP32 accumulation; twiddles,
//              small-radix passes and everything around the transform stay
//              FP32.  A wave holds 64 butterflies of a radix-16 pass as four
//              groups of 16; lane (j = lane mod 16, g = lane / 16) keeps
//              points 4g..4g+3 of butterfly j of each group, and the MFMA
//              returns outputs 4g..4g+3 to the same lane (same slot map in
//              and out, so fused kernels keep their register hand-offs).

typedef _Float16 fpm_half8 __attribute__((ext_vector_type(8)));
typedef _Float16 fpm_half2 __attribute__((ext_vector_type(2)));
typedef float fpm_float4 __attribute__((ext_vector_type(4)));
typedef float fpm_float2v __attribute__((ext_vector_type(2)));

// The forward DFT-16 as two 16 x 32 real matrices (rows: re/im of 8 outputs,
// columns: re/im of the 16 inputs), split into FP16 hi + lo, in the A-operand
// layout of v_mfma_f32_16x16x32_f16 (lane l: row l mod 16, columns
// 8 (l/16) .. +7).  Matrix h (0, 1) yields outputs 4 (p/2) + 2h + (p mod 2)
// for row pair p, so lane group g receives outputs 4g..4g+3.
struct Dft16Mats {
    fpm_half8 hi[2], lo[2];
};

__device__ __forceinline__ Dft16Mats make_dft16_mats() {
    // cos and sin of 2 pi m / 16, rounded from the exact values
    constexpr float C[16] = {1.0f, 0.92387953251128675613f, 0.70710678118654752440f, 0.38268343236508977173f,
                             0.0f, -0.38268343236508977173f, -0.70710678118654752440f, -0.92387953251128675613f,
                             -1.0f, -0.92387953251128675613f, -0.70710678118654752440f, -0.38268343236508977173f,
                             0.0f, 0.38268343236508977173f, 0.70710678118654752440f, 0.92387953251128675613f};
    const int l = threadIdx.x & 63;
    const int i = l & 15, p = i >> 1, c = i & 1;
    Dft16Mats m;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int kout = 4 * (p >> 1) + 2 * h + (p & 1);
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int kk = 8 * (l >> 4) + t;
            const int s = kk >> 1, cp = kk & 1;
            const int e = (kout * s) & 15;
            // W16^(kout s) = cos(2 pi e/16) - i sin(2 pi e/16) (forward sign)
            const float co = C[e], si = -C[(e + 12) & 15];
            // real form of y = w x: re = co xr - si xi, im = si xr + co xi
            const float a = c == 0 ? (cp == 0 ? co : -si) : (cp == 0 ? si : co);
            const _Float16 ah = (_Float16)a;
            m.hi[h][t] = ah;
            m.lo[h][t] = (_Float16)(a - (float)ah);
        }
    }
    return m;
}

// DFT-16 of the four points v[0..3] this lane holds (points 4g..4g+3 of the
// butterfly of lane j = lane mod 16); on return v[0..3] are outputs 4g..4g+3.
// Block floating point per butterfly: the inputs are scaled by 2^k so that
// the largest |component| lies in [2^14, 2^15), split exactly into
// hi = fp16(x) and lo = fp16(x - hi) (22 significant bits together), the
// product hi*hi + hi*lo + lo*hi accumulates in FP32 and the result is scaled
// back by 2^-k (exact).  The inverse is conj(DFT(conj x)).
template <bool INV> __device__ __forceinline__ void dft16_mfma(float2* v, const Dft16Mats& M) {
    float mx = 0.0f;
#pragma unroll
    for (int r = 0; r < 4; ++r) mx = fmaxf(mx, fmaxf(fabsf(v[r].x), fabsf(v[r].y)));
    mx = fmaxf(mx, __shfl_xor(mx, 16, 64));
    mx = fmaxf(mx, __shfl_xor(mx, 32, 64));
    const int k = 15 - __builtin_amdgcn_frexp_expf(mx);
    fpm_half8 bh, bl;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        fpm_float2v x;
        x.x = __builtin_amdgcn_ldexpf(v[r].x, k);
        x.y = __builtin_amdgcn_ldexpf(INV ? -v[r].y : v[r].y, k);
        const fpm_half2 h = __builtin_convertvector(x, fpm_half2);
        const fpm_float2v back = __builtin_convertvector(h, fpm_float2v);
        const fpm_half2 lo = __builtin_convertvector(x - back, fpm_half2);
        bh[2 * r] = h.x;
        bh[2 * r + 1] = h.y;
        bl[2 * r] = lo.x;
        bl[2 * r + 1] = lo.y;
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        fpm_float4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(M.lo[h], bh, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(M.hi[h], bl, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(M.hi[h], bh, acc, 0, 0, 0);
        v[2 * h].x = __builtin_amdgcn_ldexpf(acc[0], -k);
        v[2 * h].y = __builtin_amdgcn_ldexpf(INV ? -acc[1] : acc[1], -k);
        v[2 * h + 1].x = __builtin_amdgcn_ldexpf(acc[2], -k);
        v[2 * h + 1].y = __builtin_amdgcn_ldexpf(INV ? -acc[3] : acc[3], -k);
    }
}

// Engine context of a kernel: the twiddles and (split engine) the DFT-16 matrices
template <int ENG> struct EngCtx;
template <> struct EngCtx<ENG_F32> {
    __device__ __forceinline__ void init() {}
};
template <> struct EngCtx<ENG_SPLIT16> {
    Dft16Mats M;
    __device__ __forceinline__ void init() { M = make_dft16_mats(); }
};

// Slot layout of pass P: registers v[q*GS + s] (q < NG, s < GS) hold point
// first + s of butterfly q's line index j, i.e. element j + (first + s) NB.
template <int LEN, bool INV, int P, int T, class Map, int ENG> struct PassSlots {
    static constexpr int R = Plan<LEN, INV>::radix(P);
    static constexpr int NB = LEN / R;
    // the split layout needs whole waves; smaller workgroups keep FP32 butterflies
    static constexpr bool MF = ENG == ENG_SPLIT16 && R == 16 && T % 64 == 0;
    static constexpr int NG = MF ? 4 : 16 / R;
    static constexpr int GS = MF ? 4 : R;
    __device__ static __forceinline__ void group(int t, int q, int& line, int& j, int& first) {
        int b;
        if constexpr (MF) {
            b = (t >> 6) * 64 + q * 16 + (t & 15);
            first = 4 * ((t >> 4) & 3);
        } else {
            b = t + T * q;
            first = 0;
        }
        Map::split(b, NB, line, j);
    }
};

// W^(s m) for s = first .. first+3 (split engine): 3 lookups + 4 products
template <bool INV>
__device__ __forceinline__ void twiddle_quad(const Twiddles& tw, uint32_t m, int first, float2 (&w)[4]) {
    const float2 w1 = tw.get<INV>(m), w2 = tw.get<INV>(2 * m);
    const float2 wf = tw.get<INV>((uint32_t)first * m);
    w[0] = wf;
    w[1] = cmul(wf, w1);
    w[2] = cmul(wf, w2);
    w[3] = cmul(wf, cmul(w1, w2));
}

template <int R, bool INV, bool MF, int ENG>
__device__ __forceinline__ void dft_group(float2* a, const EngCtx<ENG>& ctx) {
    if constexpr (MF) dft16_mfma<INV>(a, ctx.M);
    else dft<R, INV>(a);
}

// Gather the inputs of pass P from LDS into registers, apply twiddles, DFT.
template <int LEN, bool INV, int P, int T, class Map, int ENG = ENG_F32>
__device__ __forceinline__ void pass_from_lds(float2 (&v)[16], const float2* lds, int t, const Twiddles& tw,
                                              const EngCtx<ENG>& ctx) {
    using PS = PassSlots<LEN, INV, P, T, Map, ENG>;
    constexpr int R = PS::R;
    constexpr int L = Plan<LEN, INV>::span(P);
    constexpr int NB = PS::NB;
#pragma unroll
    for (int q = 0; q < PS::NG; ++q) {
        int line, j, first;
        PS::group(t, q, line, j, first);
#pragma unroll
        for (int s = 0; s < PS::GS; ++s) v[q * PS::GS + s] = lds[Map::lds(line, j + (first + s) * NB)];
        if constexpr (L > 1) {
            const uint32_t k = (uint32_t)(j % L);
            constexpr uint32_t step = kTwiddleLen / (L * R);
            if constexpr (PS::MF) {
                float2 w[4];
                twiddle_quad<INV>(tw, k * step, first, w);
#pragma unroll
                for (int s = 0; s < 4; ++s) v[q * 4 + s] = cmul(v[q * 4 + s], w[s]);
            } else {
                float2 w[16];
                twiddle_powers<R, INV>(tw, k * step, w);
#pragma unroll
                for (int s = 1; s < R; ++s) v[q * R + s] = cmul(v[q * R + s], w[s]);
            }
        }
        dft_group<R, INV, PS::MF, ENG>(&v[q * PS::GS], ctx);
    }
}

// Scatter the outputs of pass P (held in registers) to LDS in Stockham order.
template <int LEN, bool INV, int P, int T, class Map, int ENG = ENG_F32>
__device__ __forceinline__ void pass_to_lds(const float2 (&v)[16], float2* lds, int t) {
    using PS = PassSlots<LEN, INV, P, T, Map, ENG>;
    constexpr int R = PS::R;
    constexpr int L = Plan<LEN, INV>::span(P);
#pragma unroll
    for (int q = 0; q < PS::NG; ++q) {
        int line, j, first;
        PS::group(t, q, line, j, first);
        const int base = (j / L) * L * R + (j % L);
#pragma unroll
        for (int s = 0; s < PS::GS; ++s) lds[Map::lds(line, base + (first + s) * L)] = v[q * PS::GS + s];
    }
}

// DFT of the first pass on registers loaded in the slot map of pass 0
// (pass 0 has span 1, so it has no twiddles).
template <int LEN, bool INV, int T, int ENG = ENG_F32>
__device__ __forceinline__ void first_pass_dft(float2 (&v)[16], const EngCtx<ENG>& ctx = EngCtx<ENG>{}) {
    constexpr int R = Plan<LEN, INV>::radix(0);
    constexpr bool MF = ENG == ENG_SPLIT16 && R == 16 && T % 64 == 0;
    constexpr int NG = MF ? 4 : 16 / R, GS = MF ? 4 : R;
#pragma unroll
    for (int q = 0; q < NG; ++q) dft_group<R, INV, MF, ENG>(&v[q * GS], ctx);
}

// Runs passes 1..NPASS-1 (pass 0's DFT must already be done on v).  On
// return v holds the outputs of the last pass (see Slots::out).  The caller
// must __syncthreads() before calling if the LDS was in use.
template <int LEN, bool INV, int T, class Map, int ENG = ENG_F32>
__device__ __forceinline__ void remaining_passes(float2 (&v)[16], float2* lds, int t, const Twiddles& tw,
                                                 const EngCtx<ENG>& ctx = EngCtx<ENG>{}) {
    using PL = Plan<LEN, INV>;
    static_assert(PL::NPASS <= 4, "LEN above 65536 is not supported");
    if constexpr (PL::NPASS > 1) {
        pass_to_lds<LEN, INV, 0, T, Map, ENG>(v, lds, t);
        __syncthreads();
        pass_from_lds<LEN, INV, 1, T, Map, ENG>(v, lds, t, tw, ctx);
    }
    if constexpr (PL::NPASS > 2) {
        __syncthreads();
        pass_to_lds<LEN, INV, 1, T, Map, ENG>(v, lds, t);
        __syncthreads();
        pass_from_lds<LEN, INV, 2, T, Map, ENG>(v, lds, t, tw, ctx);
    }
    if constexpr (PL::NPASS > 3) {
        __syncthreads();
        pass_to_lds<LEN, INV, 2, T, Map, ENG>(v, lds, t);
        __syncthreads();
        pass_from_lds<LEN, INV, 3, T, Map, ENG>(v, lds, t, tw, ctx);
    }
}

// Slot maps of the first pass input and the last pass output: register
// v[q*GS + s] (q < NG, s < GS) holds element (line, idx); for a fixed q the
// index steps by NB per s (idx(q, s) = idx(q, 0) + s*NB).
template <int LEN, bool INV, int T, class Map, int ENG = ENG_F32> struct Slots {
    using First = PassSlots<LEN, INV, 0, T, Map, ENG>;
    using Last = PassSlots<LEN, INV, Plan<LEN, INV>::NPASS - 1, T, Map, ENG>;
    static constexpr int R0 = First::R, NG0 = First::NG, GS0 = First::GS, NB0 = First::NB;
    static constexpr int RL = Last::R, NGL = Last::NG, GSL = Last::GS, NBL = Last::NB;
    static constexpr bool MF0 = First::MF, MFL = Last::MF;
    __device__ static __forceinline__ void in(int t, int q, int s, int& line, int& idx) {
        int j, first;
        First::group(t, q, line, j, first);
        idx = j + (first + s) * NB0;
    }
    __device__ static __forceinline__ void out(int t, int q, int s, int& line, int& idx) {
        int j, first;
        Last::group(t, q, line, j, first);
        idx = j + (first + s) * NBL;
    }
};

}  // namespace fpm
