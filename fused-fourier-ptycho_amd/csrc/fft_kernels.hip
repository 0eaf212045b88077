// fft_kernels.hip -- batched 2D C2C transforms for gfx950.
//
// Column-major R x C pages (element (i, j) at j*R + i).  A 2D transform is
//   contiguous pass : one LEN=R transform per column, a whole column in LDS;
//   strided pass(es): one LEN=C transform per row, on tiles of 16 adjacent rows
//                     (128-byte segments per column).  C <= 512 takes one
//                     pass, and so does C <= 2048 for pages in 8-row tiles
//                     (PageLayout; the tile is then one contiguous block);
//                     larger C takes two (four-step, C = M1 M2) and the
//                     frequency side is then stored column-permuted (ColPerm):
//                     forward = contiguous, A (length M1, twiddle), B (length M2)
//                     inverse = B^-1 (length M2, twiddle), A^-1 (length M1), contiguous.
// Every pass reads and writes each element once (16 B / element / pass).
// These replace the cufftPlan2d / cufftPlanMany transforms of
// cuFPM_v3/cuFPM_pure_v3.cu:118-126 (unnormalised C2C, cuFFT sign convention).
#include <algorithm>
#include <type_traits>

#include "fft_core.h"
#include "common.h"
#include "fft_kernels.h"
#include "fpm_kernels.h"

namespace fpm {

// ------------------------------------------------------- contiguous pass
constexpr int contig_threads(int len) { return len >= 16384 ? 1024 : (len >= 8192 ? 512 : 256); }
constexpr int contig_lines(int len) { return contig_threads(len) * 16 / len; }
// the thread index through an empty asm: index arithmetic that depends on it
// is formed where it is used instead of hoisted to the kernel start and kept
// (spilled) across the transform
__device__ __forceinline__ int fresh(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

template <int LEN> struct ContigCfg {
    static constexpr int T = contig_threads(LEN);
    static constexpr int LINES = contig_lines(LEN);
    using Map = LineSlowMap<LINES, LEN>;
};

template <int LEN, bool INV, int ENG>
__global__ __launch_bounds__(ContigCfg<LEN>::T) void fft_contig_kernel(const float2* __restrict__ in,
                                                                      float2* __restrict__ out,
                                                                      uint32_t nlines, Twiddles tw) {
    using Cfg = ContigCfg<LEN>;
    using Map = typename Cfg::Map;
    using SL = Slots<LEN, INV, Cfg::T, Map, ENG>;
    __shared__ float2 lds[Map::LDS_WORDS];
    __shared__ float2 twl[2 * kTwLo];
    const int t = threadIdx.x;
    tw = bind_twiddles(tw, twl, t, Cfg::T);
    EngCtx<ENG> ctx;
    ctx.init();
    const size_t line0 = (size_t)blockIdx.x * Cfg::LINES;
    float2 v[16];
#pragma unroll
    for (int q = 0; q < SL::NG0; ++q)
#pragma unroll
        for (int s = 0; s < SL::GS0; ++s) {
            int line, idx;
            SL::in(t, q, s, line, idx);
            const size_t gl = line0 + line;
            v[q * SL::GS0 + s] = gl < nlines ? in[gl * LEN + idx] : make_float2(0.f, 0.f);
        }
    first_pass_dft<LEN, INV, Cfg::T, ENG>(v, ctx);
    remaining_passes<LEN, INV, Cfg::T, Map, ENG>(v, lds, t, tw, ctx);
#pragma unroll
    for (int q = 0; q < SL::NGL; ++q)
#pragma unroll
        for (int s = 0; s < SL::GSL; ++s) {
            int line, idx;
            SL::out(t, q, s, line, idx);
            const size_t gl = line0 + line;
            if (gl < nlines) out[gl * LEN + idx] = v[q * SL::GSL + s];
        }
}

// --------------------------------------------------------- strided pass
// blockIdx.x: tile of TR rows; blockIdx.y: group g; blockIdx.z: page (x and z
// swapped for the crop source, see below).
// Point t of row r is read from column g*in_g + t*in_s and written to column
// g*out_g + t*out_s.  TWID multiplies output t by W_M^(+-g t) with
// tw_step = 16384 / M.
struct StridedArgs {
    const float2* in;
    float2* out;
    // page layout: element (row r, column c) at (r >> sh) * tstride + c * ld + (r & rmask)
    uint32_t sh, rmask, ld;
    uint64_t tstride;
    uint64_t page;  // elements per page
    uint32_t in_g, in_s, out_g, out_s;
    uint32_t tw_step;
    const uint8_t* rowmask;  // optional: skip row tiles whose 16-row entries are all 0
    // grouped tiles: page p's mask at rowmask + (p / mask_ppt) * mask_stride
    // (mask_ppt 0: one mask for every page)
    uint32_t mask_ppt, mask_stride;
    // optional: outputs stored only in natural columns c (= g cm_g + idx cm_s)
    // whose 16-column group has colmask[c / 16] != 0 (one page)
    const uint8_t* colmask;
    uint32_t cm_g, cm_s;
    // optional (one-pass forward transforms of per-LED spectra): row r stores
    // only the natural columns c whose pupil column (c + nc/2) mod nc lies in
    // the support interval pcols[(r + nr/2) mod nr] (c0 | c1 << 16) -- the
    // only points the stitch and the pupil update read; grouped tiles: page
    // p's intervals at pcols + (p / mask_ppt) * pc_nr
    const uint32_t* pcols;
    uint32_t pc_nr, pc_nc;
};

// [c0, c0 + w) of the stored columns of one row (everything without intervals)
struct StoreSpan {
    uint32_t c0 = 0, w = 0xffffffffu, cm = 0, half = 0;
    __device__ __forceinline__ void init(const StridedArgs& a, uint32_t page, uint32_t row) {
        if (!a.pcols) return;
        const uint32_t* pc = a.pcols + (a.mask_ppt ? (size_t)(page / a.mask_ppt) * a.pc_nr : 0);
        const uint32_t cc = pc[(row + (a.pc_nr >> 1)) & (a.pc_nr - 1)];
        c0 = cc & 0xffffu;
        w = (cc >> 16) - c0;  // an empty row holds 0: w = 0
        cm = a.pc_nc - 1;
        half = a.pc_nc >> 1;
    }
    __device__ __forceinline__ bool keep(uint32_t col) const { return ((col + half) & cm) - c0 < w; }
};

static void set_layout(StridedArgs& a, const PageLayout& l) {
    a.sh = l.sh;
    a.rmask = l.rmask;
    a.ld = l.ld;
    a.tstride = l.tstride;
}

// TR adjacent rows per tile: 16 (128 B per column) for long lines, more for
// short ones so that every workgroup has LEN*TR/16 >= 256 threads.
#ifndef FPM_STRIDED_POINTS
#define FPM_STRIDED_POINTS 4096  // points per strided workgroup (16 per thread)
#endif
constexpr int strided_rows_pref(int len) {
    return len >= 2048 ? 8  // 2048 x 9 x 8 B = 147 KiB of LDS
                       : len * 16 >= FPM_STRIDED_POINTS
                             ? 16
                             : (FPM_STRIDED_POINTS / len > 128 ? 128 : FPM_STRIDED_POINTS / len);
}

template <int LEN, int TR> struct StridedCfg {
    static constexpr int T = LEN * TR / 16;  // TR rows x LEN points / 16 per thread
    using Map = LineFastMap<TR, LEN>;
};

// Sources of the first pass (SRC):
//  SRC_PLAIN  a.in
//  SRC_CROP   point (row di, natural column dj) of page l is formed on the fly
//             as ifftshift(S P) of LED led0 + l -- the crop of
//             cuFPM_v3/cuda/kernels.cuh:38-64 -- gathered from the
//             column-permuted N-side spectrum, so the cropped spectrum never
//             makes a round trip to HBM;
//  SRC_STITCH point (row, natural column) of the object-spectrum gradient is
//             summed on the fly, in LED order, from the per-LED
//             back-projections R_l conj(P) / (Nr Nc) of the windows covering
//             it (the owner-computes stitch), so the gradient spectrum is
//             never written, read back or cleared.
enum { SRC_PLAIN = 0, SRC_CROP = 1, SRC_STITCH = 2 };
#ifndef FPM_STITCH_XCD
#define FPM_STITCH_XCD 1
#endif

constexpr int kMaxStitchLeds = 64;  // one bit per LED of the batch
#ifndef FPM_STITCH_PAIRS
#define FPM_STITCH_PAIRS 2  // hits of a slot whose loads the owner stitch issues together (0: one at a time)
#endif

// Crop x pupil source of one page (SRC_CROP): its tile's spectrum, pupil
// and LED (grouped tiles: page k ppt + l is LED led0 + l of tile k), and the
// pupil-support column interval [pc0, pc0 + pcw) of one pupil row (a thread's
// slots of a strided pass all lie in one row).
struct CropSrc {
    const float2* __restrict__ S;
    const float2* __restrict__ P;
    uint32_t kl, kt, pc0, pcw;
    __device__ __forceinline__ void init(const CropArgs& c, uint32_t page, uint32_t di) {
        S = c.Shat;
        P = c.P;
        const int2* cled = c.led;
        const uint32_t* cpc = c.pcols;
        uint32_t lp = page;
        if (c.tab) {
            const uint32_t tk = page / c.ppt;
            lp = page - tk * c.ppt;
            S += (size_t)tk * c.Nr * c.Nc;
            P += (size_t)tk * c.nr * c.nc;
            cled = c.tab[tk].led;
            if (cpc) cpc += (size_t)tk * c.nr;
        }
        const int2 k = cled[c.led0 + lp];
        kl = (uint32_t)k.x;
        kt = (uint32_t)k.y;
        pc0 = 0;
        pcw = 0xffffffffu;
        if (cpc) {
            const uint32_t cc = cpc[(di + (c.nr >> 1)) & (c.nr - 1)];
            pc0 = cc & 0xffffu;
            pcw = (cc >> 16) - pc0;
        }
    }
    // point (row di, natural column dj) of ifftshift(S P), cuFPM_v3/cuda/kernels.cuh:38-64
    __device__ __forceinline__ float2 at(const CropArgs& c, uint32_t di, uint32_t dj) const {
        const uint32_t i = (di + (c.nr >> 1)) & (c.nr - 1), j = (dj + (c.nc >> 1)) & (c.nc - 1);
        const uint32_t srw = (i + kt - 1u + (c.Nr >> 1)) & (c.Nr - 1);
        const uint32_t scl = (j + kl - 1u + (c.Nc >> 1)) & (c.Nc - 1);
        float2 z = make_float2(0.f, 0.f);  // P = 0: S P = +-0, S not read
        if (j - pc0 < pcw) {
            // with the support interval S is read beside P, not after it
            const float2* sp = S + (size_t)c.pN(scl) * c.Nr + srw;
            float2 sv = make_float2(0.f, 0.f);
            if (c.pcols) sv = *sp;
            const float2 pp = P[j * c.nr + i];
            if (pp.x != 0.0f || pp.y != 0.0f) {
                if (!c.pcols) sv = *sp;
                z = cmul(sv, pp);
            }
        }
        return z;
    }
    // at() in two steps, so that a loop can issue several points' loads
    // before the first product: fetch loads S beside P inside the interval
    // (without intervals: everywhere; every address is in bounds) and leaves
    // zeros outside, form gives at()'s value from the pair
    __device__ __forceinline__ void fetch(const CropArgs& c, uint32_t di, uint32_t dj, float2& sv, float2& pp) const {
        const uint32_t i = (di + (c.nr >> 1)) & (c.nr - 1), j = (dj + (c.nc >> 1)) & (c.nc - 1);
        sv = make_float2(0.f, 0.f);
        pp = make_float2(0.f, 0.f);
        if (j - pc0 < pcw) {
            const uint32_t srw = (i + kt - 1u + (c.Nr >> 1)) & (c.Nr - 1);
            const uint32_t scl = (j + kl - 1u + (c.Nc >> 1)) & (c.Nc - 1);
            sv = S[(size_t)c.pN(scl) * c.Nr + srw];
            pp = P[j * c.nr + i];
        }
    }
    __device__ __forceinline__ static float2 form(float2 sv, float2 pp) {
        return (pp.x != 0.0f || pp.y != 0.0f) ? cmul(sv, pp) : make_float2(0.f, 0.f);  // P = 0: +0, S unused
    }
};
#ifndef FPM_CROP_CHUNK
#define FPM_CROP_CHUNK 16  // points of a thread whose crop loads are issued together
#endif

// v[k] = at() of a thread's 16 input slots, gathered FPM_CROP_CHUNK points at a time
template <class SL, class Col>
__device__ __forceinline__ void crop_gather(float2 (&v)[16], const CropSrc& cs, const CropArgs& c, int tv,
                                            uint32_t row, Col col_of) {
    static_assert(SL::NG0 * SL::GS0 == 16 && 16 % FPM_CROP_CHUNK == 0, "16 input slots in whole chunks");
    constexpr int CH = FPM_CROP_CHUNK;
#pragma unroll
    for (int k0 = 0; k0 < 16; k0 += CH) {
        float2 sv[CH], pv[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            int line, idx;
            SL::in(tv, (k0 + u) / SL::GS0, (k0 + u) % SL::GS0, line, idx);
            cs.fetch(c, row, col_of(idx), sv[u], pv[u]);
        }
#pragma unroll
        for (int u = 0; u < CH; ++u) v[k0 + u] = CropSrc::form(sv[u], pv[u]);
    }
}

template <int LEN, int TR, bool INV, bool TWID, int SRC, int ENG>
__global__ __launch_bounds__(LEN * TR / 16) void fft_strided_kernel(StridedArgs a, CropArgs c, StitchArgs sa,
                                                                   Twiddles tw) {
    using Cfg = StridedCfg<LEN, TR>;
    using Map = typename Cfg::Map;
    using SL = Slots<LEN, INV, Cfg::T, Map, ENG>;
    // every slot of thread t lies in row t mod TR: true of the FP32 map
    // (butterfly t + T q) and of the split map (16 consecutive lanes walk 16
    // consecutive butterflies) when TR divides 16
    static_assert(!(SL::MF0 || SL::MFL) || 16 % TR == 0, "split engine: strided tiles of at most 16 rows");
    __shared__ float2 lds[Map::LDS_WORDS];
    __shared__ float2 twl[2 * kTwLo];
    const int t = threadIdx.x;
    // the crop launches pages fastest (grid x) so the LED pages that share a
    // pupil tile run back to back and the pupil tile stays in cache
    uint32_t tile = SRC == SRC_CROP ? blockIdx.z : blockIdx.x;
    const uint32_t page = SRC == SRC_CROP ? blockIdx.x : blockIdx.z;
    uint32_t g = blockIdx.y;
    if constexpr (SRC == SRC_STITCH && FPM_STITCH_XCD) {
        // workgroup i runs on XCD i mod 8: give each XCD a contiguous run of
        // (tile, group) pairs, so that row tiles that share the cache lines
        // at their boundaries (the windows start at any row) share an L2
        // (16384^2, batch 26: 0.81 -> 0.78 ms for the two column passes)
        const uint32_t nb = gridDim.x * gridDim.y;
        if ((nb & 7u) == 0) {
            const uint32_t b = blockIdx.x + gridDim.x * blockIdx.y;
            const uint32_t lb = (b & 7u) * (nb >> 3) + (b >> 3);
            tile = lb % gridDim.x;
            g = lb / gridDim.x;
        }
    }
    if (a.rowmask) {  // rows outside every LED window of the batch: nothing to do
        const uint8_t* rmk = a.rowmask + (a.mask_ppt ? (size_t)(page / a.mask_ppt) * a.mask_stride : 0);
        bool any = false;
        for (uint32_t k = (tile * TR) / kMaskRows; k <= (tile * TR + TR - 1) / kMaskRows; ++k)
            any |= rmk[k] != 0;
        if (!any) return;  // uniform over the workgroup
    }
    tw = bind_twiddles(tw, twl, t, Cfg::T);
    EngCtx<ENG> ctx;
    ctx.init();
    const size_t base = (size_t)page * a.page;
    const float2* __restrict__ in = a.in + base;
    float2* __restrict__ out = a.out + base;
    // LineFast: the rows of a tile are the lines, line = t mod TR for every slot
    const uint32_t row = tile * TR + (uint32_t)(t % TR);
    const size_t roff = (size_t)(row >> a.sh) * a.tstride + (row & a.rmask);
    float2 v[16];
    CropSrc cs;
    if constexpr (SRC == SRC_CROP) cs.init(c, page, row);
    StoreSpan span;
    span.init(a, page, row);
    // stitch sources (grouped tiles: page k is tile k's gradient)
    const float2* __restrict__ sR = sa.R;
    const float2* __restrict__ sP = sa.P;
    const int2* __restrict__ sled = sa.led;
    const uint32_t* __restrict__ spc = sa.pcols;
    if constexpr (SRC == SRC_STITCH) {
        if (sa.tab) {
            sR += (size_t)page * sa.b * sa.nr * sa.nc;
            sP += (size_t)page * sa.nr * sa.nc;
            sled = sa.tab[page].led;
            if (spc) spc += (size_t)page * sa.nr;
        }
    }
    // stitch: window corners of the batch, the pupil-support column interval
    // of each (LED, tile row) and this thread's row mask
    uint64_t rowhits = 0;
    uint32_t srow = 0;
    __shared__ uint32_t win[SRC == SRC_STITCH ? kMaxStitchLeds : 1];
    extern __shared__ uint32_t wcc[];  // stitch: b TR words, sized at launch
    __shared__ uint64_t colhits[SRC == SRC_STITCH ? LEN : 1];
    if constexpr (SRC == SRC_STITCH) {
        const uint32_t Rm = sa.Nr - 1, Cm = sa.Nc - 1;
        for (uint32_t l = t; l < sa.b; l += Cfg::T) {
            const int2 k = sled[sa.led0 + l];
            const uint32_t wr = ((uint32_t)k.y - 1u + (sa.Nr >> 1)) & Rm;
            const uint32_t wc = ((uint32_t)k.x - 1u + (sa.Nc >> 1)) & Cm;
            win[l] = wr | (wc << 16);
        }
        __syncthreads();
        for (uint32_t e = t; e < sa.b * TR; e += Cfg::T) {
            const uint32_t l = e / TR, li = (tile * TR + e % TR - (win[l] & 0xffffu)) & Rm;
            wcc[e] = li < sa.nr ? (spc ? spc[li] : sa.nc << 16) : 0u;  // 0: row outside the window
        }
        // the LEDs whose window covers each of the LEN column positions of the
        // workgroup's lines (position idx: column g in_g + idx in_s)
        for (uint32_t idx = t; idx < (uint32_t)LEN; idx += Cfg::T) {
            const uint32_t col = g * a.in_g + idx * a.in_s;
            uint64_t mk = 0;
            for (uint32_t l = 0; l < sa.b; ++l)
                if (((col - (win[l] >> 16)) & Cm) < sa.nc) mk |= 1ull << l;
            colhits[idx] = mk;
        }
        __syncthreads();
        srow = tile * TR + (uint32_t)(t % TR);  // LineFast: one row per thread
        for (uint32_t l = 0; l < sa.b; ++l) {
            const uint32_t cc = wcc[l * TR + t % TR];
            if ((cc >> 16) > (cc & 0xffffu)) rowhits |= 1ull << l;  // pupil rows without support add nothing
        }
    }
    if constexpr (SRC == SRC_STITCH) {
        // Each slot sums, in LED order, the LEDs whose window covers both its
        // row (rowhits) and its column (colhits of its column position): the
        // work follows the actual window hits.
        const uint32_t Rm = sa.Nr - 1, Cm = sa.Nc - 1;
        const uint32_t plane = sa.nr * sa.nc;
#pragma unroll
        for (int q = 0; q < SL::NG0; ++q)
#pragma unroll
            for (int s = 0; s < SL::GS0; ++s) {
                int line, idx;
                SL::in(t, q, s, line, idx);
                const uint32_t col = g * a.in_g + (uint32_t)idx * a.in_s;
                float2 acc = make_float2(0.f, 0.f);
                uint64_t m = rowhits & colhits[idx];
#if FPM_STITCH_PAIRS
                // the slot's hits FPM_STITCH_PAIRS at a time: their LEDs' R and P
                // loads are issued before the first product; the sum keeps LED order
                auto fetch = [&](int l, float2& rv, float2& pp) {
                    rv = make_float2(0.f, 0.f);
                    pp = make_float2(0.f, 0.f);
                    const uint32_t w = win[l];
                    const uint32_t li = (srow - (w & 0xffffu)) & Rm;
                    const uint32_t lj = (col - (w >> 16)) & Cm;  // < nc: the window covers col
                    const uint32_t cc = wcc[l * TR + t % TR];
                    const uint32_t c0 = cc & 0xffffu;
                    if (lj - c0 < (cc >> 16) - c0) {  // P = 0 outside the row's support interval
                        const uint32_t di = (li + (sa.nr >> 1)) & (sa.nr - 1);
                        const uint32_t dj = (lj + (sa.nc >> 1)) & (sa.nc - 1);
                        rv = sR[(size_t)l * plane + sa.nl.row_off(di) + (size_t)sa.nl.perm(dj) * sa.nl.ld];
                        pp = sP[lj * sa.nr + li];
                    }
                };
                auto add = [&](float2 rv, float2 pp) {
                    if (pp.x != 0.0f || pp.y != 0.0f) {  // P = 0 contributes +-0
                        const float2 gw = cmulc(rv, pp);
                        acc.x = fmaf(gw.x, sa.invN2, acc.x);
                        acc.y = fmaf(gw.y, sa.invN2, acc.y);
                    }
                };
                constexpr int K = FPM_STITCH_PAIRS;
                while (m) {
                    float2 r[K], pv[K];
#pragma unroll
                    for (int u = 0; u < K; ++u) {
                        r[u] = pv[u] = make_float2(0.f, 0.f);
                        if (m) {
                            fetch(__ffsll((unsigned long long)m) - 1, r[u], pv[u]);
                            m &= m - 1;
                        }
                    }
#pragma unroll
                    for (int u = 0; u < K; ++u) add(r[u], pv[u]);
                }
#else
                while (m) {
                    const int l = __ffsll((unsigned long long)m) - 1;
                    m &= m - 1;
                    const uint32_t w = win[l];
                    const uint32_t li = (srow - (w & 0xffffu)) & Rm;
                    const uint32_t lj = (col - (w >> 16)) & Cm;  // < nc: the window covers col
                    const uint32_t cc = wcc[l * TR + t % TR];
                    const uint32_t c0 = cc & 0xffffu;
                    if (lj - c0 < (cc >> 16) - c0) {  // P = 0 outside the row's support interval
                        const uint32_t di = (li + (sa.nr >> 1)) & (sa.nr - 1);
                        const uint32_t dj = (lj + (sa.nc >> 1)) & (sa.nc - 1);
                        const float2* rp = sR + (size_t)l * plane + sa.nl.row_off(di) + (size_t)sa.nl.perm(dj) * sa.nl.ld;
                        float2 rv = make_float2(0.f, 0.f);
                        if (spc) rv = *rp;  // R beside P inside the interval
                        const float2 pp = sP[lj * sa.nr + li];
                        if (pp.x != 0.0f || pp.y != 0.0f) {  // P = 0 contributes +-0
                            if (!spc) rv = *rp;
                            const float2 gw = cmulc(rv, pp);
                            acc.x = fmaf(gw.x, sa.invN2, acc.x);
                            acc.y = fmaf(gw.y, sa.invN2, acc.y);
                        }
                    }
                }
#endif
                v[q * SL::GS0 + s] = acc;
            }

    } else if constexpr (SRC == SRC_CROP) {
        // every slot of the thread lies in row `row` (line t mod TR)
        crop_gather<SL>(v, cs, c, t, row, [&](int idx) { return g * a.in_g + (uint32_t)idx * a.in_s; });
    } else {
#pragma unroll
        for (int q = 0; q < SL::NG0; ++q)
#pragma unroll
            for (int s = 0; s < SL::GS0; ++s) {
                int line, idx;
                SL::in(t, q, s, line, idx);
                const size_t col = (size_t)g * a.in_g + (size_t)idx * a.in_s;
                v[q * SL::GS0 + s] = in[roff + col * a.ld];
            }
    }
    first_pass_dft<LEN, INV, Cfg::T, ENG>(v, ctx);
    remaining_passes<LEN, INV, Cfg::T, Map, ENG>(v, lds, t, tw, ctx);
#pragma unroll
    for (int q = 0; q < SL::NGL; ++q)
#pragma unroll
        for (int s = 0; s < SL::GSL; ++s) {
            int line, idx;
            SL::out(t, q, s, line, idx);
            float2 r = v[q * SL::GSL + s];
            if constexpr (TWID) r = cmul(r, tw.get<INV>(g * (uint32_t)idx * a.tw_step));
            const size_t col = (size_t)g * a.out_g + (size_t)idx * a.out_s;
            if ((!a.colmask || a.colmask[(g * a.cm_g + (uint32_t)idx * a.cm_s) / kMaskRows]) && span.keep((uint32_t)col))
                out[roff + col * a.ld] = r;
        }
}

// 2048-point strided passes of 8-row tiles (the per-LED column passes at
// n = 2048) as two 512-thread workgroups per CU.  Each thread runs the slots
// of virtual threads t and t + 512 of fft_strided_kernel<2048, 8> and the
// exchanges go through half of the tile's LDS buffer, points [0, 1024) of
// the 8 rows and then [1024, 2048): 64 KiB instead of 128, so one
// workgroup's gathers and stores overlap the other's passes.  In the
// row-fast slot map a radix-16 pass's outputs of virtual thread A (t < 512)
// are the lower half of every row and B's the upper half, a radix-8 pass's
// group 0 is the lower half, and the inputs s < R/2 of every butterfly of the
// next pass lie in the lower half.  Same passes, twiddles and order as the
// one-workgroup kernel: the same bits.
#ifndef FPM_STRIDED_HALF
#define FPM_STRIDED_HALF 1
#endif
// twiddles of the first twiddled pass from an LDS table (the half-buffer kernels)
#ifndef FPM_TWIDDLE_TABLES
#define FPM_TWIDDLE_TABLES 1
#endif
namespace halfsd {
constexpr int LEN = 2048, NL = 8, TV = 1024, HALF = LEN / 2;
using FMap = LineFastMap<NL, LEN>;   // the slot arithmetic of fft_strided_kernel<2048, 8>
using HMap = LineFastMap<NL, HALF>;  // the half buffer
template <bool INV, int P> struct PassG {
    static constexpr int R = Plan<LEN, INV>::radix(P), L = Plan<LEN, INV>::span(P), NB = LEN / R, NG = 16 / R;
};
__device__ __forceinline__ int hofs(int line, int idx) { return HMap::lds(line, idx & (HALF - 1)); }

// writes of pass P's outputs of virtual thread tv, groups [q0, q1)
template <bool INV, int P>
__device__ __forceinline__ void write_groups(const float2 (&v)[16], float2* lds, int tv, int q0, int q1) {
    using G = PassG<INV, P>;
    tv = fresh(tv);
#pragma unroll
    for (int q = 0; q < G::NG; ++q) {
        if (q < q0 || q >= q1) continue;
        const int b = tv + TV * q, line = b % NL, j = b / NL;
        const int base = (j / G::L) * G::L * G::R + (j % G::L);
#pragma unroll
        for (int s = 0; s < G::R; ++s) lds[hofs(line, base + s * G::L)] = v[q * G::R + s];
    }
}
// reads of pass P's inputs s in [s0, s1) of virtual thread tv into d[q * R + s]
template <bool INV, int P>
__device__ __forceinline__ void read_points(float2 (&d)[16], const float2* lds, int tv, int s0, int s1) {
    using G = PassG<INV, P>;
    tv = fresh(tv);
#pragma unroll
    for (int q = 0; q < G::NG; ++q) {
        const int b = tv + TV * q, line = b % NL, j = b / NL;
#pragma unroll
        for (int s = 0; s < G::R; ++s)
            if (s >= s0 && s < s1) d[q * G::R + s] = lds[hofs(line, j + s * G::NB)];
    }
}
// Pass P's twiddles from an LDS table when its span is at most 16 (values
// of twiddle_powers, made once per workgroup), else from twiddle_powers.
template <bool INV, int P> constexpr bool tabled() {
    return FPM_TWIDDLE_TABLES && PassG<INV, P>::L > 1 && PassG<INV, P>::L <= 16;
}
template <bool INV, int P> constexpr int table_words() {
    return tabled<INV, P>() ? PassG<INV, P>::L * (PassG<INV, P>::R - 1) : 0;
}
template <bool INV, int P>
__device__ __forceinline__ void fill_table(float2* tab, int t, const Twiddles& tw) {
    using G = PassG<INV, P>;
    if constexpr (tabled<INV, P>()) {
        if (t < G::L) {
            constexpr uint32_t step = kTwiddleLen / (G::L * G::R);
            float2 w[16];
            twiddle_powers<G::R, INV>(tw, (uint32_t)t * step, w);
#pragma unroll
            for (int s = 1; s < G::R; ++s) tab[(s - 1) * G::L + t] = w[s];
        }
    }
}
template <bool INV, int P>
__device__ __forceinline__ void compute(float2 (&v)[16], int tv, const Twiddles& tw, const float2* tab = nullptr) {
    using G = PassG<INV, P>;
    tv = fresh(tv);
#pragma unroll
    for (int q = 0; q < G::NG; ++q) {
        if constexpr (G::L > 1) {
            const uint32_t k = (uint32_t)(((tv + TV * q) / NL) % G::L);
            if constexpr (tabled<INV, P>()) {
#pragma unroll
                for (int s = 1; s < G::R; ++s) v[q * G::R + s] = cmul(v[q * G::R + s], tab[(s - 1) * G::L + k]);
            } else {
                constexpr uint32_t step = kTwiddleLen / (G::L * G::R);
                float2 w[16];
                twiddle_powers<G::R, INV>(tw, k * step, w);
#pragma unroll
                for (int s = 1; s < G::R; ++s) v[q * G::R + s] = cmul(v[q * G::R + s], w[s]);
            }
        }
        dft<G::R, INV>(&v[q * G::R]);
    }
}
// outputs of pass P (in va / vb) -> inputs of pass P + 1 through the half
// buffer; the caller's LDS reads of any earlier exchange must be complete
template <bool INV, int P>
__device__ __forceinline__ void exchange(float2 (&va)[16], float2 (&vb)[16], float2* lds, int t) {
    using W = PassG<INV, P>;
    using Rd = PassG<INV, P + 1>;
    constexpr int H = Rd::R / 2;  // read points s < H sit in the lower half
    const int ta = t, tb = t + TV / 2;
    if constexpr (W::NG > 1) {
        // small-radix writers: groups [0, NG/2) of both virtual threads hold the lower half
        static_assert(Rd::NG == 1, "a radix-16 pass follows the small one");
        write_groups<INV, P>(va, lds, ta, 0, W::NG / 2);
        write_groups<INV, P>(vb, lds, tb, 0, W::NG / 2);
        __syncthreads();
        read_points<INV, P + 1>(va, lds, ta, 0, H);  // registers [0, H): the groups just written
        read_points<INV, P + 1>(vb, lds, tb, 0, H);
        __syncthreads();
        write_groups<INV, P>(va, lds, ta, W::NG / 2, W::NG);
        write_groups<INV, P>(vb, lds, tb, W::NG / 2, W::NG);
        __syncthreads();
        read_points<INV, P + 1>(va, lds, ta, H, Rd::R);
        read_points<INV, P + 1>(vb, lds, tb, H, Rd::R);
    } else {
        // radix-16 writers: A writes the lower half, B the upper
        float2 tmp[16];
        write_groups<INV, P>(va, lds, ta, 0, 1);
        __syncthreads();
        read_points<INV, P + 1>(va, lds, ta, 0, H);   // va is written out: free
        read_points<INV, P + 1>(tmp, lds, tb, 0, H);  // vb still holds B's outputs
        __syncthreads();
        write_groups<INV, P>(vb, lds, tb, 0, 1);
        __syncthreads();
        read_points<INV, P + 1>(va, lds, ta, H, Rd::R);
        read_points<INV, P + 1>(vb, lds, tb, H, Rd::R);
#pragma unroll
        for (int q = 0; q < Rd::NG; ++q)
#pragma unroll
            for (int s = 0; s < H; ++s) vb[q * Rd::R + s] = tmp[q * Rd::R + s];
    }
}
}  // namespace halfsd

template <bool INV, bool TWID, int SRC>
__global__ __launch_bounds__(512, 4) void fft_strided_half_kernel(StridedArgs a, CropArgs c, Twiddles tw) {
    using namespace halfsd;
    constexpr int TR = NL;
    using SL = Slots<LEN, INV, TV, FMap, ENG_F32>;
    static_assert(Plan<LEN, INV>::NPASS == 3, "2048 points: three passes");
    __shared__ float2 lds[HMap::LDS_WORDS];
    __shared__ float2 twl[2 * kTwLo];
    const int t = threadIdx.x;
    // the crop launches pages fastest (grid x), as fft_strided_kernel
    const uint32_t tile = SRC == SRC_CROP ? blockIdx.z : blockIdx.x;
    const uint32_t page = SRC == SRC_CROP ? blockIdx.x : blockIdx.z;
    const uint32_t g = blockIdx.y;
    if (a.rowmask) {  // rows outside every LED window of the batch: nothing to do
        const uint8_t* rmk = a.rowmask + (a.mask_ppt ? (size_t)(page / a.mask_ppt) * a.mask_stride : 0);
        bool any = false;
        for (uint32_t k = (tile * TR) / kMaskRows; k <= (tile * TR + TR - 1) / kMaskRows; ++k)
            any |= rmk[k] != 0;
        if (!any) return;  // uniform over the workgroup
    }
    tw = bind_twiddles(tw, twl, t, 512);
    // the table of the first twiddled pass; the exchange before its first use
    // synchronises the workgroup
    __shared__ float2 ttab[table_words<INV, 1>() > 0 ? table_words<INV, 1>() : 1];
    fill_table<INV, 1>(ttab, t, tw);
    const size_t base = (size_t)page * a.page;
    const float2* __restrict__ in = a.in + base;
    float2* __restrict__ out = a.out + base;
    // row-fast slots: every slot of virtual threads t and t + 512 lies in row t mod 8
    const uint32_t row = tile * TR + (uint32_t)(t % TR);
    const size_t roff = (size_t)(row >> a.sh) * a.tstride + (row & a.rmask);
    CropSrc cs;
    if constexpr (SRC == SRC_CROP) cs.init(c, page, row);
    StoreSpan span;
    span.init(a, page, row);
    float2 va[16], vb[16];
    auto load = [&](float2 (&v)[16], int tv) {
        tv = fresh(tv);
        if constexpr (SRC == SRC_CROP) {
            crop_gather<SL>(v, cs, c, tv, row, [&](int idx) { return g * a.in_g + (uint32_t)idx * a.in_s; });
        } else {
#pragma unroll
            for (int q = 0; q < SL::NG0; ++q)
#pragma unroll
                for (int s = 0; s < SL::GS0; ++s) {
                    int line, idx;
                    SL::in(tv, q, s, line, idx);
                    const size_t col = (size_t)g * a.in_g + (size_t)idx * a.in_s;
                    v[q * SL::GS0 + s] = in[roff + col * a.ld];
                }
        }
    };
    load(va, t);
    load(vb, t + TV / 2);
    first_pass_dft<LEN, INV, TV>(va);
    first_pass_dft<LEN, INV, TV>(vb);
    exchange<INV, 0>(va, vb, lds, t);
    compute<INV, 1>(va, t, tw, ttab);
    compute<INV, 1>(vb, t + TV / 2, tw, ttab);
    __syncthreads();
    exchange<INV, 1>(va, vb, lds, t);
    compute<INV, 2>(va, t, tw);
    compute<INV, 2>(vb, t + TV / 2, tw);
    auto store = [&](const float2 (&v)[16], int tv) {
        tv = fresh(tv);
#pragma unroll
        for (int q = 0; q < SL::NGL; ++q)
#pragma unroll
            for (int s = 0; s < SL::GSL; ++s) {
                int line, idx;
                SL::out(tv, q, s, line, idx);
                float2 r = v[q * SL::GSL + s];
                if constexpr (TWID) r = cmul(r, tw.get<INV>(g * (uint32_t)idx * a.tw_step));
                const size_t col = (size_t)g * a.out_g + (size_t)idx * a.out_s;
                if ((!a.colmask || a.colmask[(g * a.cm_g + (uint32_t)idx * a.cm_s) / kMaskRows]) &&
                    span.keep((uint32_t)col))
                    out[roff + col * a.ld] = r;
            }
    };
    store(va, t);
    store(vb, t + TV / 2);
}

// -------------------------------------------------------------- launchers
template <bool INV, int ENG>
static void launch_contig_len(int len, const float2* in, float2* out, uint32_t nlines, Twiddles tw, hipStream_t st) {
#define FPM_CONTIG_CASE(L)                                                                              \
    case L: {                                                                                           \
        dim3 grid((nlines + ContigCfg<L>::LINES - 1) / ContigCfg<L>::LINES);                            \
        hipLaunchKernelGGL((fft_contig_kernel<L, INV, ENG>), grid, dim3(ContigCfg<L>::T), 0, st, in, out, nlines, \
                           tw);                                                                         \
        break;                                                                                          \
    }
    switch (len) {
        FPM_CONTIG_CASE(16)
        FPM_CONTIG_CASE(32)
        FPM_CONTIG_CASE(64)
        FPM_CONTIG_CASE(128)
        FPM_CONTIG_CASE(256)
        FPM_CONTIG_CASE(512)
        FPM_CONTIG_CASE(1024)
        FPM_CONTIG_CASE(2048)
        FPM_CONTIG_CASE(4096)
        FPM_CONTIG_CASE(8192)
        FPM_CONTIG_CASE(16384)
        default: throw HipError{hipErrorInvalidValue, "contiguous FFT length " + std::to_string(len) + " unsupported"};
    }
#undef FPM_CONTIG_CASE
    FPM_LAUNCH_CHECK();
}

// grid.x is given in 16-row tiles; a wide-tile kernel covers TR/16 of them
struct Gather {
    const CropArgs* crop = nullptr;
    const StitchArgs* stitch = nullptr;
};

template <int L, int TR, bool INV, bool TWID, int ENG>
static void launch_strided_one(const StridedArgs& a, Gather gs, dim3 grid, Twiddles tw, hipStream_t st) {
    const CropArgs c = gs.crop ? *gs.crop : CropArgs{};
    const StitchArgs s = gs.stitch ? *gs.stitch : StitchArgs{};
    const dim3 blk(StridedCfg<L, TR>::T);
    if constexpr (FPM_STRIDED_HALF && L == halfsd::LEN && TR == halfsd::NL && ENG == ENG_F32) {
        if (!gs.stitch) {
            if constexpr (INV) {
                if (gs.crop) {
                    hipLaunchKernelGGL((fft_strided_half_kernel<INV, TWID, SRC_CROP>), dim3(grid.z, grid.y, grid.x),
                                       dim3(512), 0, st, a, c, tw);
                    return;
                }
            }
            hipLaunchKernelGGL((fft_strided_half_kernel<INV, TWID, SRC_PLAIN>), grid, dim3(512), 0, st, a, c, tw);
            return;
        }
    }
    if constexpr (INV) {
        if (gs.crop) {
            const dim3 pages_first(grid.z, grid.y, grid.x);
            hipLaunchKernelGGL((fft_strided_kernel<L, TR, INV, TWID, SRC_CROP, ENG>), pages_first, blk, 0, st, a, c, s,
                               tw);
            return;
        }
        if (gs.stitch) {
            hipLaunchKernelGGL((fft_strided_kernel<L, TR, INV, TWID, SRC_STITCH, ENG>), grid, blk,
                               (size_t)s.b * TR * sizeof(uint32_t), st, a, c, s, tw);
            return;
        }
    }
    hipLaunchKernelGGL((fft_strided_kernel<L, TR, INV, TWID, SRC_PLAIN, ENG>), grid, blk, 0, st, a, c, s, tw);
}

template <int L, bool INV, bool TWID, int ENG>
static void launch_strided_tr(const StridedArgs& a, Gather gs, dim3 grid, Twiddles tw, hipStream_t st) {
    // the split engine's slot map keeps one row per thread only for tiles of
    // at most 16 rows (the kernel asserts it)
    constexpr int TRP0 = strided_rows_pref(L);
    constexpr int TRP = (ENG == ENG_SPLIT16 && TRP0 > 16) ? 16 : TRP0;
    constexpr int TRF = TRP < 16 ? TRP : 16;  // fallback tile when rows % TRP != 0
    const uint32_t rows = grid.x * kStridedRows;
    if (rows % TRP == 0) {
        launch_strided_one<L, TRP, INV, TWID, ENG>(a, gs, dim3(rows / TRP, grid.y, grid.z), tw, st);
    } else {
        launch_strided_one<L, TRF, INV, TWID, ENG>(a, gs, dim3(rows / TRF, grid.y, grid.z), tw, st);
    }
}

template <bool INV, bool TWID, int ENG>
static void launch_strided_eng(int len, const StridedArgs& a, dim3 grid, Twiddles tw, hipStream_t st, Gather gs) {
#define FPM_STRIDED_CASE(L)                                       \
    case L:                                                       \
        launch_strided_tr<L, INV, TWID, ENG>(a, gs, grid, tw, st); \
        break;
    switch (len) {
        FPM_STRIDED_CASE(16)
        FPM_STRIDED_CASE(32)
        FPM_STRIDED_CASE(64)
        FPM_STRIDED_CASE(128)
        FPM_STRIDED_CASE(256)
        FPM_STRIDED_CASE(512)
        FPM_STRIDED_CASE(1024)
        FPM_STRIDED_CASE(2048)
        default: throw HipError{hipErrorInvalidValue, "strided FFT length " + std::to_string(len) + " unsupported"};
    }
#undef FPM_STRIDED_CASE
    FPM_LAUNCH_CHECK();
}

template <bool INV, bool TWID>
static void launch_strided_len(int len, const StridedArgs& a, dim3 grid, Twiddles tw, hipStream_t st, int eng,
                               Gather gs = Gather{}) {
    if (eng == ENG_SPLIT16) launch_strided_eng<INV, TWID, ENG_SPLIT16>(len, a, grid, tw, st, gs);
    else launch_strided_eng<INV, TWID, ENG_F32>(len, a, grid, tw, st, gs);
}

void fft_contig(const float2* in, float2* out, uint32_t len, uint32_t nlines, bool inverse, const float2* twid,
                hipStream_t st, int eng) {
    Twiddles tw{twid};
    if (eng == ENG_SPLIT16) {
        if (inverse) launch_contig_len<true, ENG_SPLIT16>((int)len, in, out, nlines, tw, st);
        else launch_contig_len<false, ENG_SPLIT16>((int)len, in, out, nlines, tw, st);
    } else {
        if (inverse) launch_contig_len<true, ENG_F32>((int)len, in, out, nlines, tw, st);
        else launch_contig_len<false, ENG_F32>((int)len, in, out, nlines, tw, st);
    }
}

// forward column transform(s): natural columns in `data` -> permuted (in
// place) or natural (into nat_out) frequency columns
void fft_cols_forward(float2* data, uint32_t rows, uint32_t cols, uint32_t batch, const ColPerm& p, const float2* twid,
                      hipStream_t st, float2* nat_out, const uint8_t* rowmask, const PageLayout* lay, int eng,
                      uint32_t mask_ppt, const uint8_t* colmask, const uint32_t* pcols) {
    if (colmask && batch != 1) throw HipError{hipErrorInvalidValue, "column mask: one page"};
    Twiddles tw{twid};
    StridedArgs a{};
    set_layout(a, lay ? *lay : colmajor_layout(rows, p));
    a.page = (uint64_t)rows * cols;
    a.rowmask = rowmask;
    a.mask_ppt = mask_ppt;
    a.mask_stride = rows / kMaskRows;
    const uint32_t tiles = rows / kStridedRows;
    if (pcols) {
        if (p.M2 != 1 || rows > 65536u || cols > 65536u)
            throw HipError{hipErrorInvalidValue, "pupil column intervals: one-pass pages of at most 65536 columns"};
        a.pcols = pcols;
        a.pc_nr = rows;
        a.pc_nc = cols;
    }
    if (p.M2 == 1) {
        a.in = data;
        a.out = nat_out ? nat_out : data;
        a.in_g = a.out_g = 0;
        a.in_s = a.out_s = 1;
        a.colmask = colmask;  // natural column idx
        a.cm_g = 0;
        a.cm_s = 1;
        launch_strided_len<false, false>((int)cols, a, dim3(tiles, 1, batch), tw, st, eng);
        return;
    }
    // pass A: length M1 over columns c2 + M2 c1, twiddle W_C^(c2 k1)
    a.in = data;
    a.out = data;
    a.in_g = a.out_g = 1;
    a.in_s = a.out_s = p.M2;
    a.tw_step = kTwiddleLen / cols;
    launch_strided_len<false, true>((int)p.M1, a, dim3(tiles, p.M2, batch), tw, st, eng);
    // pass B: length M2 over the contiguous block of group k1; output idx of
    // group k1 is natural column k1 + M1 idx
    a.in_g = p.M2;
    a.in_s = 1;
    if (nat_out) {
        a.out = nat_out;
        a.out_g = 1;
        a.out_s = p.M1;
    } else {
        a.out_g = p.M2;
        a.out_s = 1;
    }
    a.colmask = colmask;
    a.cm_g = 1;
    a.cm_s = p.M1;
    launch_strided_len<false, false>((int)p.M2, a, dim3(tiles, p.M1, batch), tw, st, eng);
}

// inverse column transform(s): permuted (or natural, from nat_in) frequency
// columns -> natural columns in `data`
void fft_cols_inverse(float2* data, uint32_t rows, uint32_t cols, uint32_t batch, const ColPerm& p, const float2* twid,
                      hipStream_t st, const float2* nat_in, const CropArgs* crop, const StitchArgs* stitch,
                      const uint8_t* rowmask, const PageLayout* lay, int eng, uint32_t mask_ppt) {
    Twiddles tw{twid};
    StridedArgs a{};
    set_layout(a, lay ? *lay : colmajor_layout(rows, p));
    a.page = (uint64_t)rows * cols;
    a.rowmask = rowmask;
    a.mask_ppt = mask_ppt;
    a.mask_stride = rows / kMaskRows;
    const uint32_t tiles = rows / kStridedRows;
    Gather gs{crop, stitch};
    if (stitch && (stitch->b > (uint32_t)kMaxStitchLeds || (batch != 1 && !stitch->tab)))
        throw HipError{hipErrorInvalidValue, "fused stitch supports one page per tile and at most 64 LEDs"};
    if (crop && crop->tab && (crop->ppt == 0 || batch % crop->ppt != 0))
        throw HipError{hipErrorInvalidValue, "grouped crop: the batch must be a multiple of the pages per tile"};
    // crop and stitch sources are addressed by natural column, like nat_in
    if (crop || stitch) nat_in = data;
    if (p.M2 == 1) {
        a.in = nat_in ? nat_in : data;
        a.out = data;
        a.in_g = a.out_g = 0;
        a.in_s = a.out_s = 1;
        launch_strided_len<true, false>((int)cols, a, dim3(tiles, 1, batch), tw, st, eng, gs);
        return;
    }
    // pass B^-1: length M2 over k2 for group k1, twiddle W_C^-(c2 k1)
    if (nat_in) {
        a.in = nat_in;
        a.in_g = 1;
        a.in_s = p.M1;
    } else {
        a.in = data;
        a.in_g = p.M2;
        a.in_s = 1;
    }
    a.out = data;
    a.out_g = p.M2;
    a.out_s = 1;
    a.tw_step = kTwiddleLen / cols;
    launch_strided_len<true, true>((int)p.M2, a, dim3(tiles, p.M1, batch), tw, st, eng, gs);
    // pass A^-1: length M1 over k1 for group c2
    a.in = data;
    a.in_g = a.out_g = 1;
    a.in_s = a.out_s = p.M2;
    launch_strided_len<true, false>((int)p.M1, a, dim3(tiles, p.M2, batch), tw, st, eng);
}

void fft2_forward(float2* data, uint32_t rows, uint32_t cols, uint32_t batch, const ColPerm& p, const float2* twid,
                  hipStream_t st, int eng) {
    fft_contig(data, data, rows, cols * batch, false, twid, st, eng);
    fft_cols_forward(data, rows, cols, batch, p, twid, st, nullptr, nullptr, nullptr, eng);
}

void fft2_inverse(float2* data, uint32_t rows, uint32_t cols, uint32_t batch, const ColPerm& p, const float2* twid,
                  hipStream_t st, int eng) {
    fft_cols_inverse(data, rows, cols, batch, p, twid, st, nullptr, nullptr, nullptr, nullptr, nullptr, eng);
    fft_contig(data, data, rows, cols * batch, true, twid, st, eng);
}


// ------------------------------------------- fused object update kernel
// Finishes the inverse transform of the object-spectrum gradient along the
// contiguous index (G holds the gradient after its column passes), applies
// RMSprop to O (cuFPM_v3/cuda/kernels.cuh:364-382: m = beta m + (1-beta)|G|^2,
// O -= lr G/(sqrt(m)+eps)) and, with FWD, starts the forward transform of the
// updated O on the same registers, writing it back into G.  The gradient never
// travels through HBM as a finished array and is never cleared: the next
// step overwrites the buffer.
// Nontemporal (streaming) loads and stores of O and m_O (1) and of the new
// spectrum (2): each is touched once per step, 5-6 GB per step at 16384^2,
// and the hint lifts the streaming rate by about 3 % on gfx950.
#ifndef FPM_NT_UPDATE
#define FPM_NT_UPDATE 2
#endif
typedef float fv2 __attribute__((ext_vector_type(2)));
template <bool NT> __device__ __forceinline__ float2 ld_s(const float2* p) {
    if constexpr (NT) {
        const fv2 r = __builtin_nontemporal_load((const fv2*)p);
        return make_float2(r.x, r.y);
    }
    return *p;
}
template <bool NT> __device__ __forceinline__ float ld_s(const float* p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    return *p;
}
template <bool NT> __device__ __forceinline__ void st_s(float2* p, float2 v) {
    if constexpr (NT) {
        fv2 r;
        r.x = v.x;
        r.y = v.y;
        __builtin_nontemporal_store(r, (fv2*)p);
    } else {
        *p = v;
    }
}
template <bool NT> __device__ __forceinline__ void st_s(float* p, float v) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// lines_per_tile > 0 (grouped tiles): the lines are several objects' back to
// back, lines_per_tile each, with one row mask per object; a workgroup's
// lines lie within one object (fused_update_groupable).
template <int LEN, bool FWD, int ENG>
__global__ __launch_bounds__(ContigCfg<LEN>::T) void fused_update_kernel(float2* __restrict__ G,
                                                                        float2* __restrict__ O,
                                                                        float* __restrict__ mO, uint32_t nlines,
                                                                        RmsArgs rms, const uint8_t* __restrict__ rowmask,
                                                                        const uint8_t* __restrict__ nextmask,
                                                                        uint32_t lines_per_tile, Twiddles tw) {
    using Cfg = ContigCfg<LEN>;
    using Map = typename Cfg::Map;
    using SI = Slots<LEN, true, Cfg::T, Map, ENG>;
    using SF = Slots<LEN, false, Cfg::T, Map, ENG>;
    static_assert(SI::RL == SF::R0, "inverse last pass and forward first pass must share the slot map");
    __shared__ float2 lds[Map::LDS_WORDS];
    __shared__ float2 twl[2 * kTwLo];
    // rows of the gradient known to be zero (outside every LED window of the
    // batch): not read, taken as 0 -- the column passes skipped them
    __shared__ uint8_t rm[LEN / kMaskRows];
    // rows of the next spectrum that the next step's column passes process
    // (its LED windows): the only rows of the forward pass's output written
    __shared__ uint8_t nm[LEN / kMaskRows];
    const int t = threadIdx.x;
    const size_t line0 = (size_t)blockIdx.x * Cfg::LINES;
    const uint8_t* rmk = rowmask;
    const uint8_t* nmk = nextmask;
    const size_t mofs = lines_per_tile ? (line0 / lines_per_tile) * (size_t)(LEN / kMaskRows) : 0;
    if (rmk) rmk += mofs;
    if (nmk) nmk += mofs;
    for (int k = t; k < LEN / kMaskRows; k += Cfg::T) {
        rm[k] = rmk ? rmk[k] : 1;
        nm[k] = nmk ? nmk[k] : 1;
    }
    tw = bind_twiddles(tw, twl, t, Cfg::T);
    EngCtx<ENG> ctx;
    ctx.init();
    const float lr = rms_lr(rms);
    float2 v[16];
#pragma unroll
    for (int q = 0; q < SI::NG0; ++q)
#pragma unroll
        for (int s = 0; s < SI::GS0; ++s) {
            int line, idx;
            SI::in(t, q, s, line, idx);
            const size_t gl = line0 + line;
            v[q * SI::GS0 + s] = (gl < nlines && rm[idx / kMaskRows]) ? G[gl * LEN + idx] : make_float2(0.f, 0.f);
        }
    first_pass_dft<LEN, true, Cfg::T, ENG>(v, ctx);
    remaining_passes<LEN, true, Cfg::T, Map, ENG>(v, lds, t, tw, ctx);
#pragma unroll
    for (int q = 0; q < SI::NGL; ++q)
#pragma unroll
        for (int s = 0; s < SI::GSL; ++s) {
            int line, idx;
            SI::out(t, q, s, line, idx);
            const size_t gl = line0 + line;
            if (gl < nlines) {
                const size_t k = gl * LEN + idx;
                const float2 gr = v[q * SI::GSL + s];
                constexpr bool NT = FPM_NT_UPDATE >= 1;
                float mu = gr.x * gr.x + gr.y * gr.y;
                mu = rms.beta * ld_s<NT>(mO + k) + (1.0f - rms.beta) * mu;
                const float sq = lr / (sqrtf(mu) + rms.eps);
                float2 w = ld_s<NT>(O + k);
                w.x -= gr.x * sq;
                w.y -= gr.y * sq;
                st_s<NT>(O + k, w);
                st_s<NT>(mO + k, mu);
                v[q * SI::GSL + s] = w;
            }
        }
    if constexpr (FWD) {
        first_pass_dft<LEN, false, Cfg::T, ENG>(v, ctx);
        __syncthreads();
        remaining_passes<LEN, false, Cfg::T, Map, ENG>(v, lds, t, tw, ctx);
#pragma unroll
        for (int q = 0; q < SF::NGL; ++q)
#pragma unroll
            for (int s = 0; s < SF::GSL; ++s) {
                int line, idx;
                SF::out(t, q, s, line, idx);
                const size_t gl = line0 + line;
                if (gl < nlines && nm[idx / kMaskRows])
                    st_s<(FPM_NT_UPDATE >= 2)>(G + gl * LEN + idx, v[q * SF::GSL + s]);
            }
    }
}

// ------------- 16384-point lines, two workgroups per CU (half-size exchange)
// A 16384-point line exchanged through LDS in one piece takes 136 KiB, so
// only one workgroup fits a CU and its load, transform and store phases
// cannot overlap another's (a contiguous pass of 16384-point lines measured
// 3.9 TB/s against 5.2 TB/s for 8192-point lines, two workgroups per CU).
// Here a 512-thread workgroup runs the 1024-thread slot maps of the
// one-piece kernel for two "virtual threads" per thread (A = t, B = t + 512)
// and exchanges through a buffer of half a line, in two phases: in every
// radix-16 write of both plans (inverse 4,16,16,16; forward 16,16,16,4)
// virtual threads below 512 write exactly the indices below 8192, and in
// the radix-4 write (inverse pass 0) groups 0-1 do; every read takes points
// s < R/2 from the lower half.  Phase 0 moves the lower half, phase 1 the
// upper, so 71 KiB of LDS per workgroup and two workgroups per CU.  Same
// slots, twiddles and butterflies as the one-piece kernel (its arithmetic,
// up to where the compiler contracts products into FMAs).  Measured at
// 16384^2, batch 26: 2.13 -> 1.88 ms per launch.
#ifndef FPM_FUSED16K_HALF
#define FPM_FUSED16K_HALF 1
#endif
#ifndef FPM_LINEAR_LDS
#define FPM_LINEAR_LDS 1
#endif
namespace half16k {
constexpr int LEN = 16384, TV = 1024, HALF = LEN / 2;
using HMap = LineSlowMap<1, HALF>;  // the half buffer (one pad word per 16)
template <bool INV, int P> struct PassG {
    static constexpr int R = Plan<LEN, INV>::radix(P), L = Plan<LEN, INV>::span(P), NB = LEN / R, NG = 16 / R;
};
__device__ __forceinline__ int hofs(int idx) { return HMap::lds(0, idx & (HALF - 1)); }
// Padded distance of d points: hofs(i + d) == hofs(i) + pad_dist(d) whenever
// i + d stays in i's half and (i mod 16) + (d mod 16) < 16.  A pass's group
// sits in one aligned block of L R <= HALF points, with base mod L < L and the
// steps multiples of L (L a power of two), and a read point j + s NB has
// j < NB, so both exchanges meet it; the addresses of one group are then one
// register plus constants that fold into the LDS instructions' offsets.
__device__ __forceinline__ constexpr int pad_dist(int d) { return d + (d >> 4); }

// writes of pass P's outputs of virtual thread tv, groups [q0, q1)
template <bool INV, int P>
__device__ __forceinline__ void write_groups(const float2 (&v)[16], float2* lds, int tv, int q0, int q1) {
    using G = PassG<INV, P>;
    tv = fresh(tv);
#pragma unroll
    for (int q = 0; q < G::NG; ++q) {
        if (q < q0 || q >= q1) continue;
        const int j = tv + TV * q;
        const int base = (j / G::L) * G::L * G::R + (j % G::L);
#if FPM_LINEAR_LDS
        float2* dst = lds + hofs(base);
#pragma unroll
        for (int s = 0; s < G::R; ++s) dst[pad_dist(s * G::L)] = v[q * G::R + s];
#else
#pragma unroll
        for (int s = 0; s < G::R; ++s) lds[hofs(base + s * G::L)] = v[q * G::R + s];
#endif
    }
}
// reads of pass P's inputs s in [s0, s1) of virtual thread tv into d[q * R + s]
template <bool INV, int P>
__device__ __forceinline__ void read_points(float2 (&d)[16], const float2* lds, int tv, int s0, int s1) {
    using G = PassG<INV, P>;
    tv = fresh(tv);
#pragma unroll
    for (int q = 0; q < G::NG; ++q) {
        const int j = tv + TV * q;
#if FPM_LINEAR_LDS
        const float2* src = lds + hofs(j);
#pragma unroll
        for (int s = 0; s < G::R; ++s)
            if (s >= s0 && s < s1) d[q * G::R + s] = src[pad_dist((s * G::NB) & (HALF - 1))];
#else
#pragma unroll
        for (int s = 0; s < G::R; ++s)
            if (s >= s0 && s < s1) d[q * G::R + s] = lds[hofs(j + s * G::NB)];
#endif
    }
}
// Pass P's twiddles W^(s k step) (k < L, s < R) from a table in LDS when its
// span L is at most 16: L (R - 1) values, made once per workgroup by
// twiddle_powers -- the same values bit for bit -- at tab[(s - 1) L + k].
template <bool INV, int P> constexpr bool tabled() {
    return FPM_TWIDDLE_TABLES && PassG<INV, P>::L > 1 && PassG<INV, P>::L <= 16;
}
template <bool INV, int P> constexpr int table_words() {
    return tabled<INV, P>() ? PassG<INV, P>::L * (PassG<INV, P>::R - 1) : 0;
}
template <bool INV, int P>
__device__ __forceinline__ void fill_table(float2* tab, int t, const Twiddles& tw) {
    using G = PassG<INV, P>;
    if constexpr (tabled<INV, P>()) {
        if (t < G::L) {
            constexpr uint32_t step = kTwiddleLen / (G::L * G::R);
            float2 w[16];
            twiddle_powers<G::R, INV>(tw, (uint32_t)t * step, w);
#pragma unroll
            for (int s = 1; s < G::R; ++s) tab[(s - 1) * G::L + t] = w[s];
        }
    }
}
// twiddles and DFT of pass P on inputs already in v (pass_from_lds without the reads)
template <bool INV, int P>
__device__ __forceinline__ void compute(float2 (&v)[16], int tv, const Twiddles& tw, const float2* tab = nullptr) {
    using G = PassG<INV, P>;
    tv = fresh(tv);
#pragma unroll
    for (int q = 0; q < G::NG; ++q) {
        if constexpr (G::L > 1) {
            const uint32_t k = (uint32_t)((tv + TV * q) % G::L);
            if constexpr (tabled<INV, P>()) {
#pragma unroll
                for (int s = 1; s < G::R; ++s) v[q * G::R + s] = cmul(v[q * G::R + s], tab[(s - 1) * G::L + k]);
            } else {
                constexpr uint32_t step = kTwiddleLen / (G::L * G::R);
                float2 w[16];
                twiddle_powers<G::R, INV>(tw, k * step, w);
#pragma unroll
                for (int s = 1; s < G::R; ++s) v[q * G::R + s] = cmul(v[q * G::R + s], w[s]);
            }
        }
        dft<G::R, INV>(&v[q * G::R]);
    }
}
// outputs of pass P (in va / vb) -> inputs of pass P + 1, through the half
// buffer; the caller's LDS reads of any earlier exchange must be complete
template <bool INV, int P>
__device__ __forceinline__ void exchange(float2 (&va)[16], float2 (&vb)[16], float2* lds, int t) {
    using W = PassG<INV, P>;
    using Rd = PassG<INV, P + 1>;
    constexpr int H = Rd::R / 2;  // read points s < H sit in the lower half
    const int ta = t, tb = t + TV / 2;
    if constexpr (W::R == 4) {
        // groups 0-1 of both virtual threads hold the lower half
        write_groups<INV, P>(va, lds, ta, 0, 2);
        write_groups<INV, P>(vb, lds, tb, 0, 2);
        __syncthreads();
        static_assert(Rd::R == 16 && Rd::NG == 1, "inverse pass 1 is radix 16");
        read_points<INV, P + 1>(va, lds, ta, 0, H);  // registers 0-7: the groups just written
        read_points<INV, P + 1>(vb, lds, tb, 0, H);
        __syncthreads();
        write_groups<INV, P>(va, lds, ta, 2, 4);
        write_groups<INV, P>(vb, lds, tb, 2, 4);
        __syncthreads();
        read_points<INV, P + 1>(va, lds, ta, H, Rd::R);
        read_points<INV, P + 1>(vb, lds, tb, H, Rd::R);
    } else {
        static_assert(W::R == 16, "radix-16 writes");
        // A (t < 512) writes the lower half, B the upper
        float2 tmp[16];
        write_groups<INV, P>(va, lds, ta, 0, 1);
        __syncthreads();
        read_points<INV, P + 1>(va, lds, ta, 0, H);   // va is written out: free
        read_points<INV, P + 1>(tmp, lds, tb, 0, H);  // vb still holds B's outputs
        __syncthreads();
        write_groups<INV, P>(vb, lds, tb, 0, 1);
        __syncthreads();
        read_points<INV, P + 1>(va, lds, ta, H, Rd::R);
        read_points<INV, P + 1>(vb, lds, tb, H, Rd::R);
#pragma unroll
        for (int q = 0; q < Rd::NG; ++q)
#pragma unroll
            for (int s = 0; s < H; ++s) vb[q * Rd::R + s] = tmp[q * Rd::R + s];
    }
}
}  // namespace half16k

template <bool FWD>
__global__ __launch_bounds__(512, 4) void fused_update16k_kernel(float2* __restrict__ G, float2* __restrict__ O,
                                                                 float* __restrict__ mO, uint32_t nlines, RmsArgs rms,
                                                                 const uint8_t* __restrict__ rowmask,
                                                                 const uint8_t* __restrict__ nextmask,
                                                                 uint32_t lines_per_tile, Twiddles tw) {
    using namespace half16k;
    using MapG = LineSlowMap<1, LEN>;  // slot arithmetic of the one-piece kernel (no LDS)
    using SI = Slots<LEN, true, TV, MapG, ENG_F32>;
    using SF = Slots<LEN, false, TV, MapG, ENG_F32>;
    __shared__ float2 lds[HMap::LDS_WORDS];
    __shared__ float2 twl[2 * kTwLo];
    // per 16-row tile: bit 0 = gradient rows read (else taken as 0), bit 1 =
    // rows of the next spectrum written (the next step's LED windows)
    __shared__ uint8_t rm[LEN / kMaskRows];
    const int t = threadIdx.x;
    const size_t line = blockIdx.x;  // one line per workgroup; the grid is nlines wide
    const uint8_t* rmk = rowmask;
    const uint8_t* nmk = nextmask;
    const size_t mofs = lines_per_tile ? (line / lines_per_tile) * (size_t)(LEN / kMaskRows) : 0;
    if (rmk) rmk += mofs;
    if (nmk) nmk += mofs;
    for (int k = t; k < LEN / kMaskRows; k += 512)
        rm[k] = (uint8_t)(((rmk ? rmk[k] : 1) != 0 ? 1 : 0) | ((nmk ? nmk[k] : 1) != 0 ? 2 : 0));
    tw = bind_twiddles(tw, twl, t, 512);
    // twiddle tables of the two small-span passes (inverse pass 1: 4 x 15,
    // forward pass 1: 16 x 15 values); the exchanges before their first use
    // synchronise the workgroup
    __shared__ float2 ttab_i[table_words<true, 1>() > 0 ? table_words<true, 1>() : 1];
    __shared__ float2 ttab_f[FWD && table_words<false, 1>() > 0 ? table_words<false, 1>() : 1];
    fill_table<true, 1>(ttab_i, t, tw);
    if constexpr (FWD) fill_table<false, 1>(ttab_f, t, tw);
    const float lr = rms_lr(rms);
    float2 va[16], vb[16];
    auto load_g = [&](float2 (&v)[16], int tv) {
#pragma unroll
        for (int q = 0; q < SI::NG0; ++q)
#pragma unroll
            for (int s = 0; s < SI::GS0; ++s) {
                int ln, idx;
                SI::in(tv, q, s, ln, idx);
                v[q * SI::GS0 + s] = (rm[idx / kMaskRows] & 1) ? G[line * LEN + idx] : make_float2(0.f, 0.f);
            }
    };
    load_g(va, t);
    load_g(vb, t + 512);
    first_pass_dft<LEN, true, TV>(va);
    first_pass_dft<LEN, true, TV>(vb);
    exchange<true, 0>(va, vb, lds, t);
    compute<true, 1>(va, t, tw, ttab_i);
    compute<true, 1>(vb, t + 512, tw, ttab_i);
    __syncthreads();
    exchange<true, 1>(va, vb, lds, t);
    compute<true, 2>(va, t, tw);
    compute<true, 2>(vb, t + 512, tw);
    __syncthreads();
    exchange<true, 2>(va, vb, lds, t);
    compute<true, 3>(va, t, tw);
    compute<true, 3>(vb, t + 512, tw);
    // RMSprop of O (cuFPM_v3/cuda/kernels.cuh:364-382) on the inverse's outputs
    auto update = [&](float2 (&v)[16], int tv) {
        tv = fresh(tv);
#pragma unroll
        for (int q = 0; q < SI::NGL; ++q)
#pragma unroll
            for (int s = 0; s < SI::GSL; ++s) {
                int ln, idx;
                SI::out(tv, q, s, ln, idx);
                const size_t k = line * LEN + idx;
                const float2 gr = v[q * SI::GSL + s];
                constexpr bool NT = FPM_NT_UPDATE >= 1;
                float mu = gr.x * gr.x + gr.y * gr.y;
                mu = rms.beta * ld_s<NT>(mO + k) + (1.0f - rms.beta) * mu;
                const float sq = lr / (sqrtf(mu) + rms.eps);
                float2 w = ld_s<NT>(O + k);
                w.x -= gr.x * sq;
                w.y -= gr.y * sq;
                st_s<NT>(O + k, w);
                st_s<NT>(mO + k, mu);
                v[q * SI::GSL + s] = w;
            }
    };
    update(va, t);
    update(vb, t + 512);
    if constexpr (FWD) {
        first_pass_dft<LEN, false, TV>(va);
        first_pass_dft<LEN, false, TV>(vb);
        __syncthreads();
        exchange<false, 0>(va, vb, lds, t);
        compute<false, 1>(va, t, tw, ttab_f);
        compute<false, 1>(vb, t + 512, tw, ttab_f);
        __syncthreads();
        exchange<false, 1>(va, vb, lds, t);
        compute<false, 2>(va, t, tw);
        compute<false, 2>(vb, t + 512, tw);
        __syncthreads();
        exchange<false, 2>(va, vb, lds, t);
        compute<false, 3>(va, t, tw);
        compute<false, 3>(vb, t + 512, tw);
        auto store = [&](const float2 (&v)[16], int tv) {
            tv = fresh(tv);
#pragma unroll
            for (int q = 0; q < SF::NGL; ++q)
#pragma unroll
                for (int s = 0; s < SF::GSL; ++s) {
                    int ln, idx;
                    SF::out(tv, q, s, ln, idx);
                    if (rm[idx / kMaskRows] & 2) st_s<(FPM_NT_UPDATE >= 2)>(G + line * LEN + idx, v[q * SF::GSL + s]);
                }
        };
        store(va, t);
        store(vb, t + 512);
    }
    (void)nlines;
}

template <int ENG>
static void launch_fused_update_eng(const Geom& g, float2* G, float2* O, float* mO, RmsArgs rms, bool forward_after,
                                    const uint8_t* rowmask, const uint8_t* nextmask, const float2* twid,
                                    hipStream_t st, uint32_t ntiles) {
    Twiddles tw{twid};
    const uint32_t nlines = g.Nc * ntiles;
    const uint32_t lpt = ntiles > 1 ? g.Nc : 0u;
    if (FPM_FUSED16K_HALF && ENG == ENG_F32 && g.Nr == 16384) {
        if (forward_after)
            hipLaunchKernelGGL((fused_update16k_kernel<true>), dim3(nlines), dim3(512), 0, st, G, O, mO, nlines, rms,
                               rowmask, nextmask, lpt, tw);
        else
            hipLaunchKernelGGL((fused_update16k_kernel<false>), dim3(nlines), dim3(512), 0, st, G, O, mO, nlines, rms,
                               rowmask, nullptr, lpt, tw);
        FPM_LAUNCH_CHECK();
        return;
    }
#define FPM_FUSED_CASE(L)                                                                                   \
    case L: {                                                                                               \
        dim3 grid((nlines + ContigCfg<L>::LINES - 1) / ContigCfg<L>::LINES);                                \
        if (forward_after)                                                                                  \
            hipLaunchKernelGGL((fused_update_kernel<L, true, ENG>), grid, dim3(ContigCfg<L>::T), 0, st, G, O, mO, nlines, \
                               rms, rowmask, nextmask, lpt, tw);                                            \
        else                                                                                                \
            hipLaunchKernelGGL((fused_update_kernel<L, false, ENG>), grid, dim3(ContigCfg<L>::T), 0, st, G, O, mO, nlines, \
                               rms, rowmask, nullptr, lpt, tw);                                             \
        break;                                                                                              \
    }
    switch (g.Nr) {
        FPM_FUSED_CASE(16)
        FPM_FUSED_CASE(32)
        FPM_FUSED_CASE(64)
        FPM_FUSED_CASE(128)
        FPM_FUSED_CASE(256)
        FPM_FUSED_CASE(512)
        FPM_FUSED_CASE(1024)
        FPM_FUSED_CASE(2048)
        FPM_FUSED_CASE(4096)
        FPM_FUSED_CASE(8192)
        FPM_FUSED_CASE(16384)
        default: throw HipError{hipErrorInvalidValue, "fused update length " + std::to_string(g.Nr) + " unsupported"};
    }
#undef FPM_FUSED_CASE
    FPM_LAUNCH_CHECK();
}

void launch_fused_update(const Geom& g, float2* G, float2* O, float* mO, RmsArgs rms, bool forward_after,
                         const uint8_t* rowmask, const uint8_t* nextmask, const float2* twid, hipStream_t st, int eng,
                         uint32_t ntiles) {
    if (ntiles > 1 && !fused_update_groupable(g))
        throw HipError{hipErrorInvalidValue, "grouped fused update: a workgroup would span two objects"};
    if (eng == ENG_SPLIT16)
        launch_fused_update_eng<ENG_SPLIT16>(g, G, O, mO, rms, forward_after, rowmask, nextmask, twid, st, ntiles);
    else
        launch_fused_update_eng<ENG_F32>(g, G, O, mO, rms, forward_after, rowmask, nextmask, twid, st, ntiles);
}

bool fused_update_groupable(const Geom& g) { return g.Nc % (uint32_t)contig_lines((int)g.Nr) == 0; }

// ------------------------ per-LED contiguous passes fused with the stencil
// One workgroup owns K = LT - 2 interior columns of one LED page plus the
// column on each side, which the stencil needs as neighbours and which are
// recomputed here.  Steps: contiguous inverse pass of every column (z stays
// in registers); e = |z|/(n_r n_c r^2) - y into LDS; the loss stencil of
// cuFPM_v3/cuda/kernels.cuh:197-275 on the interior columns, u = g z/|z|;
// the contiguous forward pass of u on the same registers (the inverse's last
// pass and the forward's first pass share the slot map).  z and u never
// reach HBM: 21 B per pixel instead of 50 for three separate kernels.
// The pages may be column-major or in 8-row tiles (Geom::nl): either way a
// wave walks 64 points down one column, 8 runs of 64 bytes in the tiled case.
#ifndef FPM_STENCIL_MAG_LDS
#define FPM_STENCIL_MAG_LDS 0
#endif
// Threads per workgroup: 1024 (16 waves, one workgroup per CU by LDS), but
// 512 at LEN 256, where 32 columns (30 interior) per workgroup cover a
// 256-column page with 288 transformed columns instead of 320 and two
// workgroups share a CU (120 VGPRs, 70 KB of LDS; the 48-tile solve with
// 24 tiles of n = 256 measured 22.07 -> 21.35 ms per epoch).  At larger LEN
// fewer threads would only add halo columns.
#ifndef FPM_STENCIL_T256
#define FPM_STENCIL_T256 512
#endif
template <int LEN> struct StencilCfg {
    static constexpr int T = LEN == 256 ? FPM_STENCIL_T256 : 1024;
    static constexpr int LT = T * 16 / LEN;  // columns per workgroup
    static constexpr int K = LT - 2;          // interior columns
    using Map = LineSlowMap<LT, LEN>;
};

// tab (grouped tiles): page k ppt + l is LED led0 + l of tile k, with tile
// k's images, normalisation and loss slot from tab[k] and its row mask at
// rowmask + k LEN/16.
// WHOLE: the workgroup holds a whole page (n_c = LT columns), every column
// is interior and its neighbours are taken cyclically from the workgroup
// (one workgroup per page instead of two with 2 interior columns in the
// second: n = 128).
// SIGNS: also write the two sign decisions of every interior pixel to
// signs (natural column-major pages: bit 0 = sign bit of er - e0, bit 1 =
// of ed - e0), a separate instantiation so the solve's kernels are unchanged.
// Residual e = |z| / ratio - y of a virtual thread's output slots (cuFPM_v3/cuda/kernels.cuh:210-222
// with the stack's amplitudes): load_y_slots issues every Y load of the slots
// together, residual_slots then converts them and forms e in LDS, and
// with_y_mode takes the uniform choices (stack type; |z| / ratio, or the exact
// product |z| (1/ratio) when ratio is a power of two) once, outside the loops,
// so that no branch between the slots holds their loads back.  The operations
// are read_y_page's and scaled_mag's.
template <class SL, bool U16, int NP, class YOff>
__device__ __forceinline__ void load_y_slots(uint32_t (&raw)[NP], int tv, const char* ypage, YOff yoff) {
    static_assert(NP == SL::NGL * SL::GSL, "one word per output slot");
#pragma unroll
    for (int q = 0; q < SL::NGL; ++q)
#pragma unroll
        for (int s = 0; s < SL::GSL; ++s) {
            int line, idx;
            SL::out(tv, q, s, line, idx);
            raw[q * SL::GSL + s] = load_y_page<U16>(ypage, yoff(line, idx));
        }
}
template <class SL, bool U16, bool QUOT, int NP, class OnMag>
__device__ __forceinline__ void residual_slots(const float2 (&v)[NP], const uint32_t (&raw)[NP], int tv, float* ef,
                                               uint32_t len, const YStack& y, float ratio, float inv_ratio,
                                               OnMag on_mag) {
#pragma unroll
    for (int q = 0; q < SL::NGL; ++q)
#pragma unroll
        for (int s = 0; s < SL::GSL; ++s) {
            int line, idx;
            SL::out(tv, q, s, line, idx);
            const float2 z = v[q * SL::GSL + s];
            const float mag = sqrtf(z.x * z.x + z.y * z.y);
            on_mag(line, idx, mag);
            const float sm = QUOT ? mag / ratio : mag * inv_ratio;
            ef[line * len + idx] = sm - y_of_raw<U16>(raw[q * SL::GSL + s], y);
        }
}
template <int LEN, int ENG, bool WHOLE, bool SIGNS>
__global__ __launch_bounds__(StencilCfg<LEN>::T, 4) void stencil_fused_kernel(const float2* __restrict__ in, float2* __restrict__ out,
                                                            Geom g, YStack y, uint32_t led0, float ratio,
                                                            float inv_ratio, float gscale,
                                                            double* __restrict__ loss_acc,
                                                            const uint8_t* __restrict__ rowmask, Twiddles tw,
                                                            const TileIn* __restrict__ tab, uint32_t ppt,
                                                            uint8_t* __restrict__ signs) {
    using Cfg = StencilCfg<LEN>;
    using Map = typename Cfg::Map;
    using SI = Slots<LEN, true, Cfg::T, Map, ENG>;
    using SF = Slots<LEN, false, Cfg::T, Map, ENG>;
    static_assert(SI::RL == SF::R0, "inverse last pass and forward first pass must share the slot map");
    static_assert(Cfg::K >= 2, "too few interior columns per workgroup");
    __shared__ float2 lds[Map::LDS_WORDS];
    __shared__ float2 twl[2 * kTwLo];
    __shared__ double red[Cfg::T / 64];
    // spectrum rows known to be zero (pupil rows without support): not read
    __shared__ uint8_t rm[LEN / kMaskRows];
    const int t = threadIdx.x;
    const uint32_t page = blockIdx.y;
    uint32_t lpage = page;
    if (tab) {
        const uint32_t tk = page / ppt;
        lpage = page - tk * ppt;
        const TileIn ti = tab[tk];
        y.data = ti.y;
        y.mn = ti.ymn;
        y.range = ti.yrange;
        y.rcp_range = ti.yrcp;
        loss_acc = ti.loss;
        if (rowmask) rowmask += (size_t)tk * (LEN / kMaskRows);
    }
    for (int k = t; k < LEN / kMaskRows; k += Cfg::T) rm[k] = rowmask ? rowmask[k] : 1;
    tw = bind_twiddles(tw, twl, t, Cfg::T);
    EngCtx<ENG> ctx;
    ctx.init();
    const uint32_t j0 = WHOLE ? 0u : blockIdx.x * Cfg::K;  // first interior column; line l is column j0 - 1 + l
    const uint32_t cmask = g.nc - 1;
    const size_t plane = (size_t)g.nr * g.nc;
    const float2* __restrict__ src = in + (size_t)page * plane;
    float2* __restrict__ dst = out + (size_t)page * plane;
    const size_t ybase = (size_t)(led0 + lpage) * plane;
    auto column = [&](int line) -> uint32_t { return WHOLE ? (uint32_t)line : (j0 - 1u + (uint32_t)line) & cmask; };
    auto interior = [&](int line) -> bool {
        return WHOLE || (line >= 1 && line <= Cfg::K && j0 + (uint32_t)line - 1u < g.nc);
    };
    // space-domain pages have natural columns in either layout
    auto at = [&](int idx, int line) -> size_t { return g.nl.row_off((uint32_t)idx) + (size_t)column(line) * g.nl.ld; };

    float2 v[16];
#pragma unroll
    for (int q = 0; q < SI::NG0; ++q)
#pragma unroll
        for (int s = 0; s < SI::GS0; ++s) {
            int line, idx;
            SI::in(t, q, s, line, idx);
            v[q * SI::GS0 + s] = rm[idx / kMaskRows] ? src[at(idx, line)] : make_float2(0.f, 0.f);
        }
    first_pass_dft<LEN, true, Cfg::T, ENG>(v, ctx);
    remaining_passes<LEN, true, Cfg::T, Map, ENG>(v, lds, t, tw, ctx);
    __syncthreads();
    // e of every column of the workgroup, in LDS (the exchange buffer as
    // floats); with FPM_STENCIL_MAG_LDS also |z| (LT*LEN words each, within
    // the LT*PADLEN float2 of the buffer), else |z| is formed again below
    float* ef = (float*)lds;
    float* mf = ef + Cfg::LT * LEN;
    static_assert(2 * Cfg::LT * LEN <= 2 * Map::LDS_WORDS, "e and |z| must fit the exchange buffer");
    // Y of this page through a global pointer and 32-bit offsets (a page
    // holds at most 2^22 points), the uniform stack-type branch outside the loop
    const char* ypage = (const char*)y.data + ybase * (y.u16 ? 2 : 4);
    int tt = fresh(t);
    auto yoff = [&](int line, int idx) { return column(line) * (uint32_t)LEN + (uint32_t)idx; };
    auto on_mag = [&](int line, int idx, float mag) {
        if constexpr (FPM_STENCIL_MAG_LDS) mf[line * LEN + idx] = mag;
    };
    if constexpr (ENG == ENG_F32) {
        with_y_mode(y, inv_ratio, [&](auto u16, auto quot) {
            constexpr bool U16 = decltype(u16)::value;
            uint32_t raw[SI::NGL * SI::GSL];
            load_y_slots<SI, U16>(raw, tt, ypage, yoff);
            residual_slots<SI, U16, decltype(quot)::value>(v, raw, tt, ef, LEN, y, ratio, inv_ratio, on_mag);
        });
    } else {
        // the split engine's register budget: one Y load per point as it is used
        auto residual = [&](auto u16) {
#pragma unroll
            for (int q = 0; q < SI::NGL; ++q)
#pragma unroll
                for (int s = 0; s < SI::GSL; ++s) {
                    int line, idx;
                    SI::out(tt, q, s, line, idx);
                    const float2 z = v[q * SI::GSL + s];
                    const float mag = sqrtf(z.x * z.x + z.y * z.y);
                    on_mag(line, idx, mag);
                    ef[line * LEN + idx] = scaled_mag(mag, ratio, inv_ratio) -
                                           read_y_page<decltype(u16)::value>(ypage, yoff(line, idx), y);
                }
        };
        if (y.u16)
            residual(std::true_type{});
        else
            residual(std::false_type{});
    }
    __syncthreads();
    float loss32 = 0.0f;  // 32 terms per thread at most; the workgroup sum is FP64
    tt = fresh(t);
#pragma unroll
    for (int q = 0; q < SI::NGL; ++q)
#pragma unroll
        for (int s = 0; s < SI::GSL; ++s) {
            int line, idx;
            SI::out(tt, q, s, line, idx);
            float2 o = make_float2(0.f, 0.f);
            if (interior(line)) {
                const float* e = ef + line * LEN;
                const float e0 = e[idx];
                const float el = WHOLE ? ef[((line + Cfg::LT - 1) & (Cfg::LT - 1)) * LEN + idx] : e[idx - LEN];
                const float er = WHOLE ? ef[((line + 1) & (Cfg::LT - 1)) * LEN + idx] : e[idx + LEN];
                const float eu = e[(idx - 1) & (LEN - 1)], ed = e[(idx + 1) & (LEN - 1)];
                const float gval = (copysignf(1.0f, e0 - el) - copysignf(1.0f, er - e0) + copysignf(1.0f, e0 - eu) -
                                    copysignf(1.0f, ed - e0)) * gscale;
                loss32 += fabsf(er - e0) + fabsf(ed - e0);
                if constexpr (SIGNS)
                    signs[(size_t)page * plane + (size_t)column(line) * LEN + idx] =
                        (uint8_t)((signbit(er - e0) ? 1 : 0) | (signbit(ed - e0) ? 2 : 0));
                const float2 z = v[q * SI::GSL + s];
                const float mag = FPM_STENCIL_MAG_LDS ? mf[line * LEN + idx] : sqrtf(z.x * z.x + z.y * z.y);
                if (mag > 0.0f) {
                    const float inv = 1.0f / mag;
                    o = make_float2(z.x * inv * gval, z.y * inv * gval);
                }
            }
            v[q * SI::GSL + s] = o;
        }
    first_pass_dft<LEN, false, Cfg::T, ENG>(v, ctx);
    __syncthreads();
    remaining_passes<LEN, false, Cfg::T, Map, ENG>(v, lds, t, tw, ctx);
    tt = fresh(t);
#pragma unroll
    for (int q = 0; q < SF::NGL; ++q)
#pragma unroll
        for (int s = 0; s < SF::GSL; ++s) {
            int line, idx;
            SF::out(tt, q, s, line, idx);
            if (interior(line) && rm[idx / kMaskRows]) dst[at(idx, line)] = v[q * SF::GSL + s];
        }
    if (loss_acc) {
        double loss = (double)loss32;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off, 64);
        if ((t & 63) == 0) red[t >> 6] = loss;
        __syncthreads();
        if (t == 0) {
            double sum = 0.0;
            for (int w = 0; w < Cfg::T / 64; ++w) sum += red[w];
            unsafeAtomicAdd(loss_acc, sum);
        }
    }
}

// -------- the fused stencil on 2048-point columns, two workgroups per CU
// stencil_fused_kernel<2048> keeps 8 columns of 2048 points in LDS (136 KiB):
// one workgroup per CU.  Here 512 threads run its 1024-thread slot maps for
// two virtual threads each (A = t, B = t + 512) and exchange through half
// the LDS, 4 columns at a time: in its radix-16 passes A owns columns 0-3
// and B columns 4-7, in its radix-8 passes group 0 of both owns columns 0-3
// and group 1 columns 4-7, so phase 0 of every exchange moves columns 0-3
// and phase 1 columns 4-7; the residuals e of all 8 columns (64 KiB of
// floats) fit the same half buffer for the stencil.  Same slots, twiddles,
// butterflies and stencil as stencil_fused_kernel<2048> (whose arithmetic
// it is, up to where the compiler contracts products into FMAs).  121
// VGPRs, no spills, 70 KiB LDS; on 26 pages of 2048^2 points 1.04 -> 0.88
// ms per launch.
#ifndef FPM_STENCIL_HALF
#define FPM_STENCIL_HALF 1
#endif
#ifndef FPM_STENCIL_SKIP_HALO
#define FPM_STENCIL_SKIP_HALO 1
#endif
namespace halfst {
constexpr int LEN = 2048, TV = 1024, NLINES = 8;
using FMap = LineSlowMap<NLINES, LEN>;  // the slot maps of stencil_fused_kernel<2048>
using HMap = LineSlowMap<NLINES / 2, LEN>;
template <bool INV, int P> struct PassG {
    static constexpr int R = Plan<LEN, INV>::radix(P), L = Plan<LEN, INV>::span(P), NB = LEN / R, NG = 16 / R;
};
__device__ __forceinline__ int hofs(int line, int idx) { return HMap::lds(line & (NLINES / 2 - 1), idx); }

template <bool INV, int P>
__device__ __forceinline__ void write_groups(const float2 (&v)[16], float2* lds, int tv, int q0, int q1) {
    using G = PassG<INV, P>;
    tv = fresh(tv);
#pragma unroll
    for (int q = 0; q < G::NG; ++q) {
        if (q < q0 || q >= q1) continue;
        const int b = tv + TV * q, line = b / G::NB, j = b % G::NB;
        const int base = (j / G::L) * G::L * G::R + (j % G::L);
#if FPM_LINEAR_LDS
        float2* dst = lds + hofs(line, base);
#pragma unroll
        for (int s = 0; s < G::R; ++s) dst[half16k::pad_dist(s * G::L)] = v[q * G::R + s];
#else
#pragma unroll
        for (int s = 0; s < G::R; ++s) lds[hofs(line, base + s * G::L)] = v[q * G::R + s];
#endif
    }
}
// inputs of pass P of groups [q0, q1), points [s0, s1), into d[q * R + s]
template <bool INV, int P>
__device__ __forceinline__ void read_groups(float2* d, const float2* lds, int tv, int q0, int q1, int s0, int s1) {
    using G = PassG<INV, P>;
    tv = fresh(tv);
#pragma unroll
    for (int q = 0; q < G::NG; ++q) {
        if (q < q0 || q >= q1) continue;
        const int b = tv + TV * q, line = b / G::NB, j = b % G::NB;
#if FPM_LINEAR_LDS
        const float2* src = lds + hofs(line, j);
#pragma unroll
        for (int s = 0; s < G::R; ++s)
            if (s >= s0 && s < s1) d[q * G::R + s] = src[half16k::pad_dist(s * G::NB)];
#else
#pragma unroll
        for (int s = 0; s < G::R; ++s)
            if (s >= s0 && s < s1) d[q * G::R + s] = lds[hofs(line, j + s * G::NB)];
#endif
    }
}
// Pass P's twiddles W^(s k step) (k < L, s < R) from a table in LDS when its
// span L is at most 16: L (R - 1) values, made once per workgroup by
// twiddle_powers -- the same values bit for bit -- at tab[(s - 1) L + k].
template <bool INV, int P> constexpr bool tabled() {
    return FPM_TWIDDLE_TABLES && PassG<INV, P>::L > 1 && PassG<INV, P>::L <= 16;
}
template <bool INV, int P> constexpr int table_words() {
    return tabled<INV, P>() ? PassG<INV, P>::L * (PassG<INV, P>::R - 1) : 0;
}
template <bool INV, int P>
__device__ __forceinline__ void fill_table(float2* tab, int t, const Twiddles& tw) {
    using G = PassG<INV, P>;
    if constexpr (tabled<INV, P>()) {
        if (t < G::L) {
            constexpr uint32_t step = kTwiddleLen / (G::L * G::R);
            float2 w[16];
            twiddle_powers<G::R, INV>(tw, (uint32_t)t * step, w);
#pragma unroll
            for (int s = 1; s < G::R; ++s) tab[(s - 1) * G::L + t] = w[s];
        }
    }
}
template <bool INV, int P>
__device__ __forceinline__ void compute(float2 (&v)[16], int tv, const Twiddles& tw, const float2* tab = nullptr) {
    using G = PassG<INV, P>;
    tv = fresh(tv);
#pragma unroll
    for (int q = 0; q < G::NG; ++q) {
        if constexpr (G::L > 1) {
            const int j = (tv + TV * q) % G::NB;
            const uint32_t k = (uint32_t)(j % G::L);
            if constexpr (tabled<INV, P>()) {
#pragma unroll
                for (int s = 1; s < G::R; ++s) v[q * G::R + s] = cmul(v[q * G::R + s], tab[(s - 1) * G::L + k]);
            } else {
                constexpr uint32_t step = kTwiddleLen / (G::L * G::R);
                float2 w[16];
                twiddle_powers<G::R, INV>(tw, k * step, w);
#pragma unroll
                for (int s = 1; s < G::R; ++s) v[q * G::R + s] = cmul(v[q * G::R + s], w[s]);
            }
        }
        dft<G::R, INV>(&v[q * G::R]);
    }
}
// outputs of pass P -> inputs of pass P + 1 through the half buffer; the LDS
// reads of any earlier use of the buffer must be complete
template <bool INV, int P>
__device__ __forceinline__ void exchange(float2 (&va)[16], float2 (&vb)[16], float2* lds, int t) {
    using W = PassG<INV, P>;
    using Rd = PassG<INV, P + 1>;
    const int ta = t, tb = t + TV / 2;
    if constexpr (W::R == 16 && Rd::R == 16) {
        // A's columns 0-3, then B's columns 4-7
        write_groups<INV, P>(va, lds, ta, 0, 1);
        __syncthreads();
        read_groups<INV, P + 1>(va, lds, ta, 0, 1, 0, 16);
        __syncthreads();
        write_groups<INV, P>(vb, lds, tb, 0, 1);
        __syncthreads();
        read_groups<INV, P + 1>(vb, lds, tb, 0, 1, 0, 16);
    } else if constexpr (W::R == 8) {
        // radix-8 writers: group 0 of both holds columns 0-3; A reads all 16
        // of its (radix-16) inputs there, 8 of them into registers whose
        // group-1 outputs are still to be written
        static_assert(Rd::R == 16, "inverse pass 1 is radix 16");
        float2 tmp[16];
        write_groups<INV, P>(va, lds, ta, 0, 1);
        write_groups<INV, P>(vb, lds, tb, 0, 1);
        __syncthreads();
        read_groups<INV, P + 1>(va, lds, ta, 0, 1, 0, 8);
        read_groups<INV, P + 1>(tmp, lds, ta, 0, 1, 8, 16);
        __syncthreads();
        write_groups<INV, P>(va, lds, ta, 1, 2);
        write_groups<INV, P>(vb, lds, tb, 1, 2);
        __syncthreads();
        read_groups<INV, P + 1>(vb, lds, tb, 0, 1, 0, 16);
#pragma unroll
        for (int s = 8; s < 16; ++s) va[s] = tmp[s];
    } else {
        // radix-16 writers, radix-8 readers: A writes columns 0-3, where group
        // 0 of both reads; B's outputs wait in vb
        static_assert(W::R == 16 && Rd::R == 8, "forward pass 2 is radix 8");
        float2 tmp[16];
        write_groups<INV, P>(va, lds, ta, 0, 1);
        __syncthreads();
        read_groups<INV, P + 1>(va, lds, ta, 0, 1, 0, 8);
        read_groups<INV, P + 1>(tmp, lds, tb, 0, 1, 0, 8);
        __syncthreads();
        write_groups<INV, P>(vb, lds, tb, 0, 1);
        __syncthreads();
        read_groups<INV, P + 1>(va, lds, ta, 1, 2, 0, 8);
        read_groups<INV, P + 1>(vb, lds, tb, 1, 2, 0, 8);
#pragma unroll
        for (int s = 0; s < 8; ++s) vb[s] = tmp[s];
    }
}
}  // namespace halfst

__global__ __launch_bounds__(512, 4) void stencil_half_kernel(const float2* __restrict__ in, float2* __restrict__ out,
                                                             Geom g, YStack y, uint32_t led0, float ratio,
                                                             float inv_ratio, float gscale,
                                                             double* __restrict__ loss_acc,
                                                             const uint8_t* __restrict__ rowmask, Twiddles tw,
                                                             const TileIn* __restrict__ tab, uint32_t ppt) {
    using namespace halfst;
    using SI = Slots<LEN, true, TV, FMap, ENG_F32>;
    using SF = Slots<LEN, false, TV, FMap, ENG_F32>;
    constexpr int K = NLINES - 2;  // interior columns
    __shared__ float2 lds[HMap::LDS_WORDS];
    __shared__ float2 twl[2 * kTwLo];
    __shared__ double red[512 / 64];
    __shared__ uint8_t rm[LEN / kMaskRows];
    const int t = threadIdx.x;
    const uint32_t page = blockIdx.y;
    uint32_t lpage = page;
    if (tab) {
        const uint32_t tk = page / ppt;
        lpage = page - tk * ppt;
        const TileIn ti = tab[tk];
        y.data = ti.y;
        y.mn = ti.ymn;
        y.range = ti.yrange;
        y.rcp_range = ti.yrcp;
        loss_acc = ti.loss;
        if (rowmask) rowmask += (size_t)tk * (LEN / kMaskRows);
    }
    for (int k = t; k < LEN / kMaskRows; k += 512) rm[k] = rowmask ? rowmask[k] : 1;
    tw = bind_twiddles(tw, twl, t, 512);
    // twiddle tables of the two small-span passes (inverse pass 1: 8 x 15,
    // forward pass 1: 16 x 15 values); the exchanges before their first use
    // synchronise the workgroup
    __shared__ float2 ttab_i[table_words<true, 1>() > 0 ? table_words<true, 1>() : 1];
    __shared__ float2 ttab_f[table_words<false, 1>() > 0 ? table_words<false, 1>() : 1];
    fill_table<true, 1>(ttab_i, t, tw);
    fill_table<false, 1>(ttab_f, t, tw);
    const uint32_t j0 = blockIdx.x * K;  // first interior column; line l is column j0 - 1 + l
    const uint32_t cmask = g.nc - 1;
    const size_t plane = (size_t)g.nr * g.nc;
    const float2* __restrict__ src = in + (size_t)page * plane;
    float2* __restrict__ dst = out + (size_t)page * plane;
    const size_t ybase = (size_t)(led0 + lpage) * plane;
    auto column = [&](int line) -> uint32_t { return (j0 - 1u + (uint32_t)line) & cmask; };
    auto interior = [&](int line) -> bool { return line >= 1 && line <= K && j0 + (uint32_t)line - 1u < g.nc; };
    auto at = [&](int idx, int line) -> size_t { return g.nl.row_off((uint32_t)idx) + (size_t)column(line) * g.nl.ld; };

    float2 va[16], vb[16];
    auto load = [&](float2 (&v)[16], int tv) {
        tv = fresh(tv);
#pragma unroll
        for (int q = 0; q < SI::NG0; ++q)
#pragma unroll
            for (int s = 0; s < SI::GS0; ++s) {
                int line, idx;
                SI::in(tv, q, s, line, idx);
                v[q * SI::GS0 + s] = rm[idx / kMaskRows] ? src[at(idx, line)] : make_float2(0.f, 0.f);
            }
    };
    load(va, t);
    load(vb, t + 512);
    first_pass_dft<LEN, true, TV>(va);
    first_pass_dft<LEN, true, TV>(vb);
    exchange<true, 0>(va, vb, lds, t);
    compute<true, 1>(va, t, tw, ttab_i);
    compute<true, 1>(vb, t + 512, tw, ttab_i);
    __syncthreads();
    exchange<true, 1>(va, vb, lds, t);
    compute<true, 2>(va, t, tw);
    compute<true, 2>(vb, t + 512, tw);
    __syncthreads();
    // e of every column of the workgroup, in LDS (the half buffer as floats)
    float* ef = (float*)lds;
    static_assert(NLINES * LEN <= 2 * HMap::LDS_WORDS, "e must fit the half buffer");
    // Y of this page through a global pointer and 32-bit offsets (a page
    // holds 2^22 points), the uniform stack-type branch outside the loop
    const char* ypage = (const char*)y.data + ybase * (y.u16 ? 2 : 4);
    auto yoff = [&](int line, int idx) { return column(line) * (uint32_t)LEN + (uint32_t)idx; };
    auto no_mag = [](int, int, float) {};
    with_y_mode(y, inv_ratio, [&](auto u16, auto quot) {
        constexpr bool U16 = decltype(u16)::value, Q = decltype(quot)::value;
        uint32_t ra[SI::NGL * SI::GSL], rb[SI::NGL * SI::GSL];
        // both virtual threads' loads before either residual: 32 in flight
        load_y_slots<SI, U16>(ra, fresh(t), ypage, yoff);
        load_y_slots<SI, U16>(rb, fresh(t + 512), ypage, yoff);
        residual_slots<SI, U16, Q>(va, ra, fresh(t), ef, LEN, y, ratio, inv_ratio, no_mag);
        residual_slots<SI, U16, Q>(vb, rb, fresh(t + 512), ef, LEN, y, ratio, inv_ratio, no_mag);
    });
    __syncthreads();
    // FP32 partial sums of 32 terms per virtual thread, as in the 1024-thread kernel
    float loss_a = 0.0f, loss_b = 0.0f;
    auto stencil = [&](float2 (&v)[16], int tv, float& loss32) {
        tv = fresh(tv);
#pragma unroll
        for (int q = 0; q < SI::NGL; ++q)
#pragma unroll
            for (int s = 0; s < SI::GSL; ++s) {
                int line, idx;
                SI::out(tv, q, s, line, idx);
                float2 o = make_float2(0.f, 0.f);
                if (interior(line)) {
                    const float* e = ef + line * LEN;
                    const float e0 = e[idx];
                    const float el = e[idx - LEN], er = e[idx + LEN];
                    const float eu = e[(idx - 1) & (LEN - 1)], ed = e[(idx + 1) & (LEN - 1)];
                    const float gval = (copysignf(1.0f, e0 - el) - copysignf(1.0f, er - e0) +
                                        copysignf(1.0f, e0 - eu) - copysignf(1.0f, ed - e0)) * gscale;
                    loss32 += fabsf(er - e0) + fabsf(ed - e0);
                    const float2 z = v[q * SI::GSL + s];
                    const float mag = sqrtf(z.x * z.x + z.y * z.y);
                    if (mag > 0.0f) {
                        const float inv = 1.0f / mag;
                        o = make_float2(z.x * inv * gval, z.y * inv * gval);
                    }
                }
                v[q * SI::GSL + s] = o;
            }
    };
    stencil(va, t, loss_a);
    stencil(vb, t + 512, loss_b);
#if FPM_STENCIL_SKIP_HALO
    // The forward plan's first two passes (radix 16, 128 butterflies per
    // line) keep virtual thread tv on line tv / 128: the halo lines' inputs
    // are zero and their outputs are not stored, so the waves of a halo line
    // skip them (wave-uniform; every line is its own transform, so the other
    // lines' bits do not change).  Each SIMD holds one such half-idle wave.
    static_assert(SI::NGL == 1 && SI::GSL == 16 && PassG<false, 0>::R == 16 && PassG<false, 1>::R == 16,
                  "forward passes 0 and 1 are radix 16 on the inverse's output layout");
    const bool run_a = interior(t / (LEN / 16)), run_b = interior((t + 512) / (LEN / 16));
#else
    const bool run_a = true, run_b = true;
#endif
    if (run_a) first_pass_dft<LEN, false, TV>(va);
    if (run_b) first_pass_dft<LEN, false, TV>(vb);
    __syncthreads();
    exchange<false, 0>(va, vb, lds, t);
    if (run_a) compute<false, 1>(va, t, tw, ttab_f);
    if (run_b) compute<false, 1>(vb, t + 512, tw, ttab_f);
    __syncthreads();
    exchange<false, 1>(va, vb, lds, t);
    compute<false, 2>(va, t, tw);
    compute<false, 2>(vb, t + 512, tw);
    auto store = [&](const float2 (&v)[16], int tv) {
        tv = fresh(tv);
#pragma unroll
        for (int q = 0; q < SF::NGL; ++q)
#pragma unroll
            for (int s = 0; s < SF::GSL; ++s) {
                int line, idx;
                SF::out(tv, q, s, line, idx);
                if (interior(line) && rm[idx / kMaskRows]) dst[at(idx, line)] = v[q * SF::GSL + s];
            }
    };
    store(va, t);
    store(vb, t + 512);
    if (loss_acc) {
        double loss = (double)loss_a + (double)loss_b;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off, 64);
        if ((t & 63) == 0) red[t >> 6] = loss;
        __syncthreads();
        if (t == 0) {
            double sum = 0.0;
            for (int w = 0; w < 512 / 64; ++w) sum += red[w];
            unsafeAtomicAdd(loss_acc, sum);
        }
    }
}

bool stencil_fused_supported(const Geom& g) { return g.nr <= 2048; }

// ------------------------------------ whole per-LED pipeline of small pages
// One workgroup runs every per-LED stage of a 128 x 128 page: the crop x
// pupil gathered from the object spectrum (kernels.cuh:19-65, as the
// SRC_CROP source of fft_strided_kernel forms it), the inverse transform
// along the columns index, the inverse along the rows index, the residual
// and loss stencil with cyclic neighbours (kernels.cuh:197-275, as the
// whole-page stencil_fused_kernel<128> computes them), the forward
// transform along the rows index and the forward along the columns index;
// only R = DFT_n(u) of the rows with pupil support is written.  The page
// (128 KiB) stays in LDS between the directions: a transform along the
// columns index uses LineFastMap<128, 128> (lines = rows), one along the
// rows index LineSlowMap<128, 128> (lines = columns), and each change of
// direction is one LDS write in the one map and one read in the other.  It
// replaces three launches (crop + inverse column pass, fused stencil,
// forward column pass) and their two HBM round trips of the pages.
// PERSIST: a grid of at most one workgroup per CU walks the pages
// (page = blockIdx.x, blockIdx.x + gridDim.x, ...).
constexpr int kPageLen = 128;
#ifndef FPM_PAGE_FUSED
#define FPM_PAGE_FUSED 1
#endif
#ifndef FPM_PAGE_PERSIST
#define FPM_PAGE_PERSIST 0
#endif

template <int ENG, bool PERSIST>
__global__ __launch_bounds__(1024, 4) void page_fused_kernel(CropArgs c, float2* __restrict__ out, Geom g, YStack y0,
                                                           float ratio, float inv_ratio, float gscale,
                                                           double* __restrict__ loss0,
                                                           const uint8_t* __restrict__ rowmask0, Twiddles tw,
                                                           const TileIn* __restrict__ tab, uint32_t npages) {
    constexpr int LEN = kPageLen, T = 1024;
    using FMap = LineFastMap<LEN, LEN>;
    using SMap = LineSlowMap<LEN, LEN>;
    using SJI = Slots<LEN, true, T, FMap, ENG>;   // inverse along the columns index
    using SCI = Slots<LEN, true, T, SMap, ENG>;   // inverse along the rows index
    using SCF = Slots<LEN, false, T, SMap, ENG>;  // forward along the rows index
    using SJF = Slots<LEN, false, T, FMap, ENG>;  // forward along the columns index
    static_assert(SCI::RL == SCF::R0, "inverse last pass and forward first pass must share the slot map");
    constexpr int WORDS = FMap::LDS_WORDS > SMap::LDS_WORDS ? FMap::LDS_WORDS : SMap::LDS_WORDS;
    __shared__ float2 lds[WORDS];
    __shared__ float2 twl[2 * kTwLo];
    __shared__ double red[T / 64];
    __shared__ uint8_t rm[LEN / kMaskRows];
    const int t = threadIdx.x;
    tw = bind_twiddles(tw, twl, t, T);
    EngCtx<ENG> ctx;
    ctx.init();
    constexpr uint32_t plane = (uint32_t)LEN * LEN;
    constexpr uint32_t h = LEN / 2, m = LEN - 1;
    for (uint32_t page = blockIdx.x; page < npages; page += PERSIST ? gridDim.x : npages) {
        // this page's tile (grouped tiles: page k ppt + l is LED led0 + l of tile k)
        uint32_t lp = page;
        const float2* __restrict__ cS = c.Shat;
        const float2* __restrict__ cP = c.P;
        const int2* cled = c.led;
        YStack y = y0;
        double* loss_acc = loss0;
        const uint8_t* rowmask = rowmask0;
        if (tab) {
            const uint32_t tk = page / c.ppt;
            lp = page - tk * c.ppt;
            cS += (size_t)tk * c.Nr * c.Nc;
            cP += (size_t)tk * plane;
            const TileIn ti = tab[tk];
            cled = ti.led;
            y.data = ti.y;
            y.mn = ti.ymn;
            y.range = ti.yrange;
            y.rcp_range = ti.yrcp;
            loss_acc = ti.loss;
            if (rowmask) rowmask += (size_t)tk * (LEN / kMaskRows);
        }
        if (PERSIST) __syncthreads();  // the previous page's last reads of rm and red
        for (int k = t; k < LEN / kMaskRows; k += T) rm[k] = rowmask ? rowmask[k] : 1;
        const int2 kk = cled[c.led0 + lp];
        const uint32_t kl = (uint32_t)kk.x, kt = (uint32_t)kk.y;
        // crop x pupil at (row di, natural column dj): ifftshift(S P)
        float2 v[16];
        int tt = fresh(t);
        if constexpr (ENG == ENG_F32) {
            // every P load of the thread, then the S loads where P != 0, then
            // the products: two round trips instead of one per point
            float2 pv[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                int line, idx;
                SJI::in(tt, k / SJI::GS0, k % SJI::GS0, line, idx);
                const uint32_t i = ((uint32_t)line + h) & m, j = ((uint32_t)idx + h) & m;
                pv[k] = cP[j * LEN + i];
            }
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                int line, idx;
                SJI::in(tt, k / SJI::GS0, k % SJI::GS0, line, idx);
                const uint32_t i = ((uint32_t)line + h) & m, j = ((uint32_t)idx + h) & m;
                const uint32_t srw = (i + kt - 1u + (c.Nr >> 1)) & (c.Nr - 1);
                const uint32_t scl = (j + kl - 1u + (c.Nc >> 1)) & (c.Nc - 1);
                v[k] = make_float2(0.f, 0.f);
                if (pv[k].x != 0.0f || pv[k].y != 0.0f) v[k] = cS[(size_t)c.pN(scl) * c.Nr + srw];
            }
#pragma unroll
            for (int k = 0; k < 16; ++k)  // P = 0: S P = +-0, S not read
                v[k] = (pv[k].x != 0.0f || pv[k].y != 0.0f) ? cmul(v[k], pv[k]) : make_float2(0.f, 0.f);
        } else {
#pragma unroll
        for (int q = 0; q < SJI::NG0; ++q)
#pragma unroll
            for (int s = 0; s < SJI::GS0; ++s) {
                int line, idx;
                SJI::in(tt, q, s, line, idx);
                const uint32_t i = ((uint32_t)line + h) & m, j = ((uint32_t)idx + h) & m;
                const uint32_t srw = (i + kt - 1u + (c.Nr >> 1)) & (c.Nr - 1);
                const uint32_t scl = (j + kl - 1u + (c.Nc >> 1)) & (c.Nc - 1);
                const float2 pp = cP[j * LEN + i];
                float2 z = make_float2(0.f, 0.f);  // P = 0: S P = +-0, S not read
                if (pp.x != 0.0f || pp.y != 0.0f) z = cmul(cS[(size_t)c.pN(scl) * c.Nr + srw], pp);
                v[q * SJI::GS0 + s] = z;
            }
        }
        first_pass_dft<LEN, true, T, ENG>(v, ctx);
        if (PERSIST) __syncthreads();  // the previous page's last reads of lds
        remaining_passes<LEN, true, T, FMap, ENG>(v, lds, fresh(t), tw, ctx);
        // columns index done: on to the rows index through LDS
        __syncthreads();
        tt = fresh(t);
#pragma unroll
        for (int q = 0; q < SJI::NGL; ++q)
#pragma unroll
            for (int s = 0; s < SJI::GSL; ++s) {
                int line, idx;
                SJI::out(tt, q, s, line, idx);
                lds[SMap::lds(idx, line)] = v[q * SJI::GSL + s];
            }
        __syncthreads();
        tt = fresh(t);
#pragma unroll
        for (int q = 0; q < SCI::NG0; ++q)
#pragma unroll
            for (int s = 0; s < SCI::GS0; ++s) {
                int line, idx;
                SCI::in(tt, q, s, line, idx);
                v[q * SCI::GS0 + s] = lds[SMap::lds(line, idx)];
            }
        first_pass_dft<LEN, true, T, ENG>(v, ctx);
        __syncthreads();
        remaining_passes<LEN, true, T, SMap, ENG>(v, lds, fresh(t), tw, ctx);
        __syncthreads();
        // residual e of every column into LDS (floats), then the stencil
        float* ef = (float*)lds;
        const char* ypage = (const char*)y.data + (size_t)(c.led0 + lp) * plane * (y.u16 ? 2 : 4);
        tt = fresh(t);
        with_y_mode(y, inv_ratio, [&](auto u16, auto quot) {
            constexpr bool U16 = decltype(u16)::value;
            uint32_t raw[SCI::NGL * SCI::GSL];
            load_y_slots<SCI, U16>(raw, tt, ypage,
                                   [&](int line, int idx) { return (uint32_t)line * LEN + (uint32_t)idx; });
            residual_slots<SCI, U16, decltype(quot)::value>(v, raw, tt, ef, LEN, y, ratio, inv_ratio,
                                                            [](int, int, float) {});
        });
        __syncthreads();
        float loss32 = 0.0f;
        tt = fresh(t);
#pragma unroll
        for (int q = 0; q < SCI::NGL; ++q)
#pragma unroll
            for (int s = 0; s < SCI::GSL; ++s) {
                int line, idx;
                SCI::out(tt, q, s, line, idx);
                const float* e = ef + line * LEN;
                const float e0 = e[idx];
                const float el = ef[((line + LEN - 1) & m) * LEN + idx];
                const float er = ef[((line + 1) & m) * LEN + idx];
                const float eu = e[(idx - 1) & m], ed = e[(idx + 1) & m];
                const float gval = (copysignf(1.0f, e0 - el) - copysignf(1.0f, er - e0) + copysignf(1.0f, e0 - eu) -
                                    copysignf(1.0f, ed - e0)) * gscale;
                loss32 += fabsf(er - e0) + fabsf(ed - e0);
                const float2 z = v[q * SCI::GSL + s];
                const float mag = sqrtf(z.x * z.x + z.y * z.y);
                float2 o = make_float2(0.f, 0.f);
                if (mag > 0.0f) {
                    const float inv = 1.0f / mag;
                    o = make_float2(z.x * inv * gval, z.y * inv * gval);
                }
                v[q * SCI::GSL + s] = o;
            }
        first_pass_dft<LEN, false, T, ENG>(v, ctx);
        __syncthreads();
        remaining_passes<LEN, false, T, SMap, ENG>(v, lds, fresh(t), tw, ctx);
        // rows index done: on to the columns index through LDS
        __syncthreads();
        tt = fresh(t);
#pragma unroll
        for (int q = 0; q < SCF::NGL; ++q)
#pragma unroll
            for (int s = 0; s < SCF::GSL; ++s) {
                int line, idx;
                SCF::out(tt, q, s, line, idx);
                lds[FMap::lds(idx, line)] = v[q * SCF::GSL + s];
            }
        __syncthreads();
        tt = fresh(t);
#pragma unroll
        for (int q = 0; q < SJF::NG0; ++q)
#pragma unroll
            for (int s = 0; s < SJF::GS0; ++s) {
                int line, idx;
                SJF::in(tt, q, s, line, idx);
                v[q * SJF::GS0 + s] = lds[FMap::lds(line, idx)];
            }
        first_pass_dft<LEN, false, T, ENG>(v, ctx);
        __syncthreads();
        remaining_passes<LEN, false, T, FMap, ENG>(v, lds, fresh(t), tw, ctx);
        float2* __restrict__ dst = out + (size_t)page * plane;
        tt = fresh(t);
#pragma unroll
        for (int q = 0; q < SJF::NGL; ++q)
#pragma unroll
            for (int s = 0; s < SJF::GSL; ++s) {
                int line, idx;
                SJF::out(tt, q, s, line, idx);
                if (rm[line / kMaskRows]) dst[g.nl.off((uint32_t)line, (uint32_t)idx)] = v[q * SJF::GSL + s];
            }
        if (loss_acc) {
            double loss = (double)loss32;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off, 64);
            if ((t & 63) == 0) red[t >> 6] = loss;
            __syncthreads();
            if (t == 0) {
                double sum = 0.0;
                for (int w = 0; w < T / 64; ++w) sum += red[w];
                unsafeAtomicAdd(loss_acc, sum);
            }
        }
    }
}

bool page_fused_supported(const Geom& g) {
    return FPM_PAGE_FUSED && !g.generic && g.nr == (uint32_t)kPageLen && g.nc == (uint32_t)kPageLen;
}

void launch_page_fused(const Geom& g, const CropArgs& ca, float2* rec, const YStack& y, uint32_t pages, int pratio,
                       double* loss_acc, const uint8_t* rowmask, const float2* twid, hipStream_t st, int eng) {
    if (!page_fused_supported(g)) throw HipError{hipErrorInvalidValue, "fused page pipeline: pages of 128 x 128 only"};
    if (ca.tab && (ca.ppt == 0 || pages % ca.ppt != 0))
        throw HipError{hipErrorInvalidValue, "fused page pipeline: the pages must be a multiple of the pages per tile"};
    if (pages == 0) return;
    Twiddles tw{twid};
    const float ratio = (float)(g.plane() * (uint32_t)(pratio * pratio));
    const float gscale = (float)(pratio * pratio);
    const float inv_ratio = exact_inverse_or_zero(ratio);
    uint32_t grid = pages;
    if (FPM_PAGE_PERSIST) {
        int dev = 0, cus = 0;
        FPM_HIP(hipGetDevice(&dev));
        FPM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        grid = std::min<uint32_t>(pages, (uint32_t)std::max(cus, 1));
    }
#define FPM_PF_LAUNCH(E)                                                                                         \
    hipLaunchKernelGGL((page_fused_kernel<E, FPM_PAGE_PERSIST != 0>), dim3(grid), dim3(1024), 0, st, ca, rec, g, y, \
                       ratio, inv_ratio, gscale, loss_acc, rowmask, tw, ca.tab, pages)
    if (eng == ENG_SPLIT16)
        FPM_PF_LAUNCH(ENG_SPLIT16);
    else
        FPM_PF_LAUNCH(ENG_F32);
#undef FPM_PF_LAUNCH
    FPM_LAUNCH_CHECK();
}

template <int ENG, bool SIGNS>
static void launch_stencil_fused_eng(const Geom& g, const float2* z_cols, const YStack& y, uint32_t led0, uint32_t b,
                                     int pratio, float2* u_out, double* loss_acc, const uint8_t* rowmask,
                                     const float2* twid, hipStream_t st, const TileIn* tab, uint32_t ppt,
                                     uint8_t* signs) {
    Twiddles tw{twid};
    const float ratio = (float)(g.plane() * (uint32_t)(pratio * pratio));
    const float gscale = (float)(pratio * pratio);
    const float inv_ratio = exact_inverse_or_zero(ratio);
    if (FPM_STENCIL_HALF && ENG == ENG_F32 && !SIGNS && g.nr == 2048 && g.nc != (uint32_t)StencilCfg<2048>::LT) {
        dim3 grid((g.nc + 5) / 6, b);
        hipLaunchKernelGGL(stencil_half_kernel, grid, dim3(512), 0, st, z_cols, u_out, g, y, led0, ratio, inv_ratio,
                           gscale, loss_acc, rowmask, tw, tab, ppt);
        FPM_LAUNCH_CHECK();
        return;
    }
#define FPM_STF_CASE(L)                                                                                   \
    case L: {                                                                                             \
        if (g.nc == (uint32_t)StencilCfg<L>::LT) {                                                        \
            hipLaunchKernelGGL((stencil_fused_kernel<L, ENG, true, SIGNS>), dim3(1, b), dim3(StencilCfg<L>::T), 0, st, \
                               z_cols, u_out, g, y, led0, ratio, inv_ratio, gscale, loss_acc, rowmask, tw, tab, ppt, \
                               signs);                                                                    \
        } else {                                                                                          \
            dim3 grid((g.nc + StencilCfg<L>::K - 1) / StencilCfg<L>::K, b);                               \
            hipLaunchKernelGGL((stencil_fused_kernel<L, ENG, false, SIGNS>), grid, dim3(StencilCfg<L>::T), 0, st, \
                               z_cols, u_out, g, y, led0, ratio, inv_ratio, gscale, loss_acc, rowmask, tw, tab, ppt, \
                               signs);                                                                    \
        }                                                                                                 \
        break;                                                                                            \
    }
    switch (g.nr) {
        FPM_STF_CASE(16)
        FPM_STF_CASE(32)
        FPM_STF_CASE(64)
        FPM_STF_CASE(128)
        FPM_STF_CASE(256)
        FPM_STF_CASE(512)
        FPM_STF_CASE(1024)
        FPM_STF_CASE(2048)
        default: throw HipError{hipErrorInvalidValue, "fused stencil length " + std::to_string(g.nr) + " unsupported"};
    }
#undef FPM_STF_CASE
    FPM_LAUNCH_CHECK();
}

void launch_stencil_fused(const Geom& g, const float2* z_cols, const YStack& y, uint32_t led0, uint32_t b, int pratio,
                          float2* u_out, double* loss_acc, const uint8_t* rowmask, const float2* twid,
                          hipStream_t st, int eng, const TileIn* tab, uint32_t ppt, uint8_t* signs) {
    if (tab && (ppt == 0 || b % ppt != 0))
        throw HipError{hipErrorInvalidValue, "grouped stencil: the batch must be a multiple of the pages per tile"};
    if (tab && signs) throw HipError{hipErrorInvalidValue, "fused stencil: sign output is not available for tiles"};
    if (b > 65535) throw HipError{hipErrorInvalidValue, "fused stencil: at most 65535 pages per launch"};
    if (eng == ENG_SPLIT16) {
        if (signs)
            launch_stencil_fused_eng<ENG_SPLIT16, true>(g, z_cols, y, led0, b, pratio, u_out, loss_acc, rowmask, twid,
                                                        st, tab, ppt, signs);
        else
            launch_stencil_fused_eng<ENG_SPLIT16, false>(g, z_cols, y, led0, b, pratio, u_out, loss_acc, rowmask, twid,
                                                         st, tab, ppt, nullptr);
    } else {
        if (signs)
            launch_stencil_fused_eng<ENG_F32, true>(g, z_cols, y, led0, b, pratio, u_out, loss_acc, rowmask, twid, st,
                                                    tab, ppt, signs);
        else
            launch_stencil_fused_eng<ENG_F32, false>(g, z_cols, y, led0, b, pratio, u_out, loss_acc, rowmask, twid, st,
                                                     tab, ppt, nullptr);
    }
}

}  // namespace fpm
