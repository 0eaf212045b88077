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

namespace fpm {

// ------------------------------------------------------------------ complex
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
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

// exp(-+ 2 pi i m / 8) applied to a (m in 1..3)
template <bool INV, int M> __device__ __forceinline__ float2 w8(float2 a) {
    constexpr float h = 0.70710678118654752440f;
    if constexpr (M == 1) {  // (1 - i)/sqrt2 forward
        return INV ? make_float2(h * (a.x - a.y), h * (a.x + a.y)) : make_float2(h * (a.x + a.y), h * (a.y - a.x));
    } else if constexpr (M == 2) {
        return mul_mi<INV>(a);
    } else {  // M == 3: (-1 - i)/sqrt2 forward
        return INV ? make_float2(-h * (a.x + a.y), h * (a.x - a.y)) : make_float2(h * (a.y - a.x), -h * (a.x + a.y));
    }
}

// radix-8 as 4 x 2: n = 2 n1 + n2, k = k1 + 4 k2
template <bool INV> __device__ __forceinline__ void dft8(float2* a) {
    // DFT4 over n1 for each n2
    dft4<INV>(a[0], a[2], a[4], a[6]);
    dft4<INV>(a[1], a[3], a[5], a[7]);
    // twiddle W8^(n2 k1), n2 = 1
    a[3] = w8<INV, 1>(a[3]);
    a[5] = w8<INV, 2>(a[5]);
    a[7] = w8<INV, 3>(a[7]);
    // DFT2 over n2 for each k1
    dft2<INV>(a[0], a[1]);
    dft2<INV>(a[2], a[3]);
    dft2<INV>(a[4], a[5]);
    dft2<INV>(a[6], a[7]);
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
template <int NLINES, int LEN> struct LineFastMap {
    static constexpr int LDS_WORDS = LEN * (NLINES + 1);
    __device__ static __forceinline__ void split(int b, int NB, int& line, int& j) { line = b % NLINES; j = b / NLINES; }
    __device__ static __forceinline__ int lds(int line, int idx) { return idx * (NLINES + 1) + line; }
};

// Gather the inputs of pass P from LDS into registers, apply twiddles, DFT.
template <int LEN, bool INV, int P, int T, class Map>
__device__ __forceinline__ void pass_from_lds(float2 (&v)[16], const float2* lds, int t, const Twiddles& tw) {
    using PL = Plan<LEN, INV>;
    constexpr int R = PL::radix(P);
    constexpr int L = PL::span(P);
    constexpr int NB = LEN / R;  // butterflies per line
#pragma unroll
    for (int q = 0; q < 16 / R; ++q) {
        int line, j;
        Map::split(t + T * q, NB, line, j);
#pragma unroll
        for (int s = 0; s < R; ++s) v[q * R + s] = lds[Map::lds(line, j + s * NB)];
        if constexpr (L > 1) {
            const uint32_t k = (uint32_t)(j % L);
            constexpr uint32_t step = kTwiddleLen / (L * R);
            float2 w[16];
            twiddle_powers<R, INV>(tw, k * step, w);
#pragma unroll
            for (int s = 1; s < R; ++s) v[q * R + s] = cmul(v[q * R + s], w[s]);
        }
        dft<R, INV>(&v[q * R]);
    }
}

// Scatter the outputs of pass P (held in registers) to LDS in Stockham order.
template <int LEN, bool INV, int P, int T, class Map>
__device__ __forceinline__ void pass_to_lds(const float2 (&v)[16], float2* lds, int t) {
    using PL = Plan<LEN, INV>;
    constexpr int R = PL::radix(P);
    constexpr int L = PL::span(P);
    constexpr int NB = LEN / R;
#pragma unroll
    for (int q = 0; q < 16 / R; ++q) {
        int line, j;
        Map::split(t + T * q, NB, line, j);
        const int base = (j / L) * L * R + (j % L);
#pragma unroll
        for (int s = 0; s < R; ++s) lds[Map::lds(line, base + s * L)] = v[q * R + s];
    }
}

// DFT of the first pass on registers loaded as v[q*R + s] = x[line][j + s*LEN/R]
// (pass 0 has span 1, so it has no twiddles).
template <int LEN, bool INV>
__device__ __forceinline__ void first_pass_dft(float2 (&v)[16]) {
    constexpr int R = Plan<LEN, INV>::radix(0);
#pragma unroll
    for (int q = 0; q < 16 / R; ++q) dft<R, INV>(&v[q * R]);
}

// Runs passes 1..NPASS-1 (pass 0's DFT must already be done on v).  On
// return v holds the outputs of the last pass (see Slots::out).  The caller
// must __syncthreads() before calling if the LDS was in use.
template <int LEN, bool INV, int T, class Map>
__device__ __forceinline__ void remaining_passes(float2 (&v)[16], float2* lds, int t, const Twiddles& tw) {
    using PL = Plan<LEN, INV>;
    static_assert(PL::NPASS <= 4, "LEN above 65536 is not supported");
    if constexpr (PL::NPASS > 1) {
        pass_to_lds<LEN, INV, 0, T, Map>(v, lds, t);
        __syncthreads();
        pass_from_lds<LEN, INV, 1, T, Map>(v, lds, t, tw);
    }
    if constexpr (PL::NPASS > 2) {
        __syncthreads();
        pass_to_lds<LEN, INV, 1, T, Map>(v, lds, t);
        __syncthreads();
        pass_from_lds<LEN, INV, 2, T, Map>(v, lds, t, tw);
    }
    if constexpr (PL::NPASS > 3) {
        __syncthreads();
        pass_to_lds<LEN, INV, 2, T, Map>(v, lds, t);
        __syncthreads();
        pass_from_lds<LEN, INV, 3, T, Map>(v, lds, t, tw);
    }
}

// slot (q, s) of thread t -> (line, index) for the first pass input and the
// last pass output
template <int LEN, bool INV, int T, class Map> struct Slots {
    static constexpr int R0 = Plan<LEN, INV>::radix(0);
    static constexpr int RL = Plan<LEN, INV>::radix(Plan<LEN, INV>::NPASS - 1);
    __device__ static __forceinline__ void in(int t, int q, int s, int& line, int& idx) {
        constexpr int NB = LEN / R0;
        int j;
        Map::split(t + T * q, NB, line, j);
        idx = j + s * NB;
    }
    __device__ static __forceinline__ void out(int t, int q, int s, int& line, int& idx) {
        constexpr int NB = LEN / RL;
        int j;
        Map::split(t + T * q, NB, line, j);
        idx = j + s * NB;
    }
};

}  // namespace fpm
