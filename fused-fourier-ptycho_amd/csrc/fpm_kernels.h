// fpm_kernels.h -- host launchers of the FPM hot-path kernels.
#pragma once
#include "common.h"
#include <type_traits>

namespace fpm {

// Geometry of one solve: object N_rows x N_cols, pupil/images n_rows x n_cols.
struct Geom {
    uint32_t nr, nc;   // low-res rows / cols
    uint32_t Nr, Nc;   // high-res rows / cols
    uint32_t lnr;      // log2 nr
    ColPerm pn;        // column permutation of n-side spectra (column-major pages)
    ColPerm pN;        // column permutation of the N-side spectrum
    PageLayout nl;     // layout of the per-LED pages of a minibatch (tiled or column-major with pn)
    int eng = ENG_F32; // FFT engine of every transform of the solve
    bool generic = false;  // a side is not a power of two: the generic (mixed-radix, unfused) path
    uint32_t plane() const { return nr * nc; }
};

// image stack reader: float amplitudes, or uint16 raw counts with
// y = sqrt((I - mn) / (mx - mn)) in float
struct YStack {
    const void* data;
    int32_t u16;       // 0: float amplitudes, 1: uint16 raw
    uint32_t mn;       // raw minimum (u16 only)
    float range;       // (float)(mx - mn) (u16 only)
    float rcp_range;   // 1.0f / range rounded to nearest (u16 only)
};

// a / b rounded to nearest for integers 0 <= a < 2^16, 1 <= b < 2^16, given
// rb = RN(1/b): q = RN(a rb) is within an ulp, the remainder a - q b is
// exact under FMA and one correction q + r rb rounds correctly (Markstein's
// division theorem).  Three operations instead of the IEEE division
// sequence; equality with a / b for every such (a, b) is checked
// exhaustively by the test suite.
__device__ __forceinline__ float div_u16(float a, float b, float rb) {
    const float q = a * rb;
    const float r = fmaf(-q, b, a);
    return fmaf(r, rb, q);
}

// amplitude at element `off` of the stack (uint16: (I - min)/(max - min) in
// float, then sqrt -- the single-precision steps of main_run_FPM.m:43-45)
__device__ __forceinline__ float read_y(const YStack& y, size_t off) {
    if (y.u16) {
        const uint32_t raw = ((const uint16_t*)y.data)[off];
        return sqrtf(div_u16((float)(raw - y.mn), y.range, y.rcp_range));
    }
    return ((const float*)y.data)[off];
}

// read_y at element `off` of one page (`page` = the page's first element,
// the stack type fixed at compile time): loads through a global-space
// pointer and a 32-bit offset, so a page's loads need no 64-bit addresses
template <bool U16>
__device__ __forceinline__ uint32_t load_y_page(const char* page, uint32_t off) {
    if constexpr (U16) return ((const __attribute__((address_space(1))) uint16_t*)page)[off];
    else return __float_as_uint(((const __attribute__((address_space(1))) float*)page)[off]);
}
// the amplitude of a raw word of load_y_page (split from the load, so that a
// loop can issue all of its loads before the first conversion)
template <bool U16>
__device__ __forceinline__ float y_of_raw(uint32_t raw, const YStack& y) {
    if constexpr (U16) return sqrtf(div_u16((float)(raw - y.mn), y.range, y.rcp_range));
    else return __uint_as_float(raw);
}
template <bool U16>
__device__ __forceinline__ float read_y_page(const char* page, uint32_t off, const YStack& y) {
    return y_of_raw<U16>(load_y_page<U16>(page, off), y);
}

// |z| / ratio as the reference's absC (cuFPM_v3/cuda/kernels.cuh:5-9); when
// ratio = n_r n_c r^2 is a power of two, inv_ratio = 1/ratio is exact and the
// product equals the quotient; otherwise inv_ratio is 0 and the division runs
__device__ __forceinline__ float scaled_mag(float mag, float ratio, float inv_ratio) {
    return inv_ratio > 0.0f ? mag * inv_ratio : mag / ratio;
}

// f(U16, QUOT) for the uniform choices of a residual loop, taken once outside
// it: the stack type, and scaled_mag's quotient (QUOT) or exact product
template <class F>
__device__ __forceinline__ void with_y_mode(const YStack& y, float inv_ratio, F f) {
    const bool quot = !(inv_ratio > 0.0f);  // scaled_mag's test
    if (y.u16) {
        if (quot) f(std::true_type{}, std::true_type{});
        else f(std::true_type{}, std::false_type{});
    } else {
        if (quot) f(std::false_type{}, std::true_type{});
        else f(std::false_type{}, std::false_type{});
    }
}

// the raw word of read_y at element `off` (uint16 count or float bits); y_of_raw converts it
template <bool U16>
__device__ __forceinline__ uint32_t load_y_raw(const YStack& y, size_t off) {
    if constexpr (U16) return ((const uint16_t*)y.data)[off];
    else return __float_as_uint(((const float*)y.data)[off]);
}

// RMSprop + pupil constraint constants of one step.  lr_dev (optional): the
// learning rate lives in device memory (the solve keeps it there so that the
// epoch-end kernel can halve it and every epoch's launches are identical,
// which lets an epoch be captured once as a graph and replayed); else lr.
struct RmsArgs {
    float lr, beta, eps, clamp;
    const float* lr_dev = nullptr;
};
__device__ __forceinline__ float rms_lr(const RmsArgs& r) { return r.lr_dev ? *r.lr_dev : r.lr; }

// solve state in device memory: learning rate, epoch counter, FP64 loss of
// the running epoch and the per-epoch loss history
struct SolveState {
    float* lr;
    uint32_t* epoch;
    double* loss_cur;   // may be null (no loss requested)
    double* loss_hist;  // [epochs], may be null
    uint32_t halve_from = 13;  // lr halves after every epoch with index >= halve_from
    // grouped tiles: loss_cur[k] and loss_hist[k * hist_stride + epoch] per tile k
    uint32_t ntiles = 1, hist_stride = 0;
};
// lr = lr0, epoch = 0, loss_cur = 0 (a kernel, not a memset: graph-safe)
void launch_solve_state_init(const SolveState& s, float lr0, hipStream_t st);
// end of an epoch: loss_hist[epoch] = loss_cur, loss_cur = 0, lr halved after
// every epoch with index >= halve_from (13: cuFPM_v3/cuFPM_pure_v3.cu:204-206),
// epoch += 1
void launch_epoch_end(const SolveState& s, hipStream_t st);

// feature-domain loss gradient (kernels.cuh:197-275): z (b pages, natural) and
// images led0.. -> u.  loss_acc (device double, may be null) += sum |Dx e|+|Dy e|.
// signs (optional): one byte per pixel of the b pages (page l pixel (i, j) at
// l n_r n_c + j n_r + i), bit 0 = sign bit of e(i, j+1) - e(i, j), bit 1 = of
// e(i+1, j) - e(i, j): the decisions the stencil took
void launch_stencil(const Geom& g, const float2* z, const YStack& y, uint32_t led0, uint32_t b, int pratio,
                    float2* u, double* loss_acc, hipStream_t st, uint8_t* signs = nullptr);

// back-projection (kernels.cuh:67-129): pupil gradient R conj(S) summed in
// registers over the batch in LED order (S read back from the spectrum Shat
// at the window element) and, in kBackprojAtomic mode, FP32 atomics of
// R conj(P) / (Nr Nc) into Gs (kBackprojPupil: the owner-computes stitch is
// done elsewhere).  If mP and A are non-null the pupil RMSprop + clamp of
// kernels.cuh:384-410 is applied to P in place; otherwise the pupil gradient
// is written to GP.
enum { kBackprojPupil = 0, kBackprojAtomic = 1 };
// tab (optional, ntiles entries): grouped tiles, tile k = grid row k with
// R pages k b.., Shat, P, Gs and mP blocks k, LED windows and A from tab[k]
void launch_backproj(const Geom& g, const float2* R, const float2* Shat, const int2* led, uint32_t led0, uint32_t b,
                     float2* P, float2* Gs, int mode, float* mP, const float* A, RmsArgs rms, float2* GP,
                     hipStream_t st, const TileIn* tab = nullptr, uint32_t ntiles = 1);

// owner-computes stitch: every element of Gs written as the sum, in LED order,
// of R_l . conj(P) / (Nr Nc) over the windows covering it (zero elsewhere).
// tab (optional): grouped tiles, grid z = tile k (R pages k b.., P and Gs
// blocks k, LED windows from tab[k])
// pcols (optional): the pupil-support column interval of each pupil row
// (launch_pupil_row_mask; grouped tiles: nr words per tile), P read only inside
void launch_owner_stitch(const Geom& g, const float2* R, const float2* P, const int2* led, uint32_t led0, uint32_t b,
                         float2* Gs, hipStream_t st, const TileIn* tab = nullptr, uint32_t ntiles = 1,
                         const uint32_t* pcols = nullptr);

// contiguous inverse pass + loss stencil + contiguous forward pass of a batch
// of n-side pages in one kernel (z_cols: pages after their inverse column
// passes; u_out: pages before their forward column passes); nr <= 2048.
// rowmask (optional, nr/16 bytes): input rows of 0-entries are taken as zero and
// output rows of 0-entries are not written (the forward column pass skips them).
// tab (optional): grouped tiles, b = ntiles x ppt pages, tile k = pages
// k ppt..; its images, normalisation, loss slot from tab[k], its row mask at
// rowmask + k nr/16.
bool stencil_fused_supported(const Geom& g);
// signs (optional, not with tab): the stencil's decisions as launch_stencil writes them
void launch_stencil_fused(const Geom& g, const float2* z_cols, const YStack& y, uint32_t led0, uint32_t b, int pratio,
                          float2* u_out, double* loss_acc, const uint8_t* rowmask, const float2* twid,
                          hipStream_t st, int eng = ENG_F32, const TileIn* tab = nullptr, uint32_t ppt = 0,
                          uint8_t* signs = nullptr);

// the whole per-LED pipeline of 128 x 128 pages in one kernel: crop x pupil
// of ca's spectrum, both inverse passes, the residual and loss stencil, both
// forward passes; rec receives R (rows of rowmask's nonzero 16-row entries;
// all rows without a mask) in g.nl's layout.  ca.tab / ca.ppt: grouped
// tiles (pages = ntiles x ppt; images, normalisation, loss slot and row mask
// of tile k as for launch_stencil_fused).
struct CropArgs;
bool page_fused_supported(const Geom& g);
void launch_page_fused(const Geom& g, const CropArgs& ca, float2* rec, const YStack& y, uint32_t pages, int pratio,
                       double* loss_acc, const uint8_t* rowmask, const float2* twid, hipStream_t st, int eng);

// contiguous inverse pass of the object gradient + RMSprop of O (kernels.cuh:364-382)
// [+ contiguous forward pass of the new O, written back into G]
// rowmask (optional, Nr/16 bytes): gradient rows of 0-entries are taken as zero
// nextmask (optional, Nr/16 bytes): only rows of the forward pass's output in
// 16-row tiles of nonzero entries are written -- the rows the next step's
// column passes process (its LED windows); the others are left as they were
// ntiles > 1: grouped tiles, G, O, mO hold ntiles objects back to back and
// rowmask / nextmask ntiles masks (fused_update_groupable must hold)
void launch_fused_update(const Geom& g, float2* G, float2* O, float* mO, RmsArgs rms, bool forward_after,
                         const uint8_t* rowmask, const uint8_t* nextmask, const float2* twid, hipStream_t st,
                         int eng = ENG_F32, uint32_t ntiles = 1);
// every workgroup of the fused update stays within one object's lines
bool fused_update_groupable(const Geom& g);

// per 16-row tile of the N-side spectrum: 1 if a LED window of the batch
// covers a row of the tile at a pupil row i with sup[i] != 0 (sup: nr bytes,
// or nullptr for every pupil row); Nr / 16 bytes
// tab (optional): grouped tiles, tile k's LED windows from tab[k], its
// support at sup + k nr and its mask at mask + k Nr/16.
// axis 1: the same for 16-column groups of the spectrum (Nr, nr, sup the
// column count, pupil columns and their support; the LEDs' kl offsets)
void launch_row_mask(const int2* led, uint32_t led0, uint32_t b, uint32_t Nr, uint32_t nr, const uint8_t* sup,
                     uint8_t* mask, hipStream_t st, const TileIn* tab = nullptr, uint32_t ntiles = 1,
                     int axis = 0);
// csup (nc bytes) = 1 for the pupil columns inside the support interval of
// some pupil row (cols: nr words from launch_pupil_row_mask)
void launch_pupil_col_support(const uint32_t* cols, uint32_t nr, uint32_t nc, uint8_t* csup, hipStream_t st);

// pupil support: sup[i] (nr bytes) = 1 iff pupil row i holds P != 0 or A != 0;
// mask (nr/16 bytes): the same per 16-row tile of the ifftshifted per-LED
// spectra -- the only rows where S P can be nonzero and where R is read.
// cols (optional, nr words): cols[i] = c0 | c1 << 16 with [c0, c1) the
// smallest column interval of pupil row i holding every P != 0 and A != 0
// (0 when the row has none), so P = 0 outside it at every step.
// Valid for every step of a solve with the clamped update.
void launch_pupil_row_mask(const float2* P, const float* A, uint32_t nr, uint32_t nc, uint8_t* mask, uint8_t* sup,
                           hipStream_t st, uint32_t* cols = nullptr);

// fill with zeros (count complex elements)
void launch_zero(float2* p, size_t count, hipStream_t st);

// |z| / (nr nc) (or its square) of b natural pages -> float pages
void launch_abs_scaled(const float2* z, size_t count, float scale, int squared, float* out, hipStream_t st);

// uint16 stack min/max (device) and conversion to float amplitudes
void launch_u16_minmax(const uint16_t* raw, size_t count, uint32_t* minmax /* device [2] */, hipStream_t st);
void launch_u16_to_amp(const uint16_t* raw, size_t count, uint32_t mn, float range, float* out, hipStream_t st);

}  // namespace fpm
