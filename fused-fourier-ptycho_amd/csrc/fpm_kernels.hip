// fpm_kernels.hip -- FPM hot-path kernels for gfx950 (crop x pupil, loss
// stencil, back-projection + stitch, pupil update, helpers).
//
// Layouts: caller arrays (object, pupil, images) are column-major, pixel
// (i, j) at j*rows + i.  The object spectrum is column-major and stored
// column-permuted by the FFT's ColPerm; the per-LED pages of a minibatch use
// Geom::nl (8-row tiles, natural columns, when both sides are <= 2048;
// column-major with ColPerm pn otherwise).  The fftshift of cuFPM_v3 is folded
// into the index maps (k + M/2) mod M, valid for even M
// (cuFPM_v3/cuda/kernels.cuh:45-47,62-64,92-93,109-111).
#include "fft_core.h"
#include "common.h"
#include "fpm_kernels.h"
#include <algorithm>

namespace fpm {

__device__ __forceinline__ float sgnf(float x) { return copysignf(1.0f, x); }

// ------------------------------------------------------------ loss stencil
// cuFPM_v3/cuda/kernels.cuh:197-275 (differenceKernel_acc):
//   e = |z|/(n_r n_c r^2) - y,  sx = sgn(e(i,j+1) - e(i,j)),  sy = sgn(e(i+1,j) - e(i,j))
//   g = r^2 (sx(i,j-1) - sx(i,j) + sy(i-1,j) - sy(i,j)),  u = g z/|z|  (0 where |z| = 0)
// all circular.  A workgroup owns a 64-row x 32-column tile; e of the tile and
// a one-pixel halo is staged in LDS once, so every z and y is read once (plus
// the halo) and the row-contiguous accesses are coalesced.
constexpr int kStTR = 64, kStTC = 32, kStLD = kStTR + 2;

__global__ __launch_bounds__(256) void stencil_kernel(Geom g, const float2* __restrict__ z, YStack y, uint32_t led0,
                                                      float ratio, float inv_ratio, float gscale,
                                                      float2* __restrict__ u, double* __restrict__ loss_acc,
                                                      uint8_t* __restrict__ signs) {
    __shared__ float es[(kStTC + 2) * kStLD];
    __shared__ double red[4];
    const int t = threadIdx.x;
    const int r = t & 63, cg = t >> 6;
    const uint32_t row0 = blockIdx.x * kStTR, col0 = blockIdx.y * kStTC;
    const uint32_t ell = blockIdx.z;
    const uint32_t plane = g.nr * g.nc;
    const size_t page = (size_t)ell * plane;
    const size_t ypage = (size_t)(led0 + ell) * plane;
    const uint32_t rmask = g.nr - 1, cmask = g.nc - 1;

    auto e_at = [&](uint32_t gi, uint32_t gj, float2* zout) -> float {
        const uint32_t off = gj * g.nr + gi;
        const float2 zz = z[page + off];
        if (zout) *zout = zz;
        return scaled_mag(sqrtf(zz.x * zz.x + zz.y * zz.y), ratio, inv_ratio) - read_y(y, ypage + off);
    };

    float2 zc[8];
    const uint32_t gi = (row0 + r) & rmask;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int c = cg * 8 + k;
        es[(c + 1) * kStLD + r + 1] = e_at(gi, (col0 + c) & cmask, &zc[k]);
    }
    if (t < 128) {  // left / right halo columns
        const int c = t < 64 ? -1 : kStTC;
        es[(c + 1) * kStLD + r + 1] = e_at(gi, (col0 + c) & cmask, nullptr);
    } else if (t < 192) {  // top / bottom halo rows
        const int q = t - 128;
        const int hr = q < 32 ? -1 : kStTR;
        const int c = q & 31;
        es[(c + 1) * kStLD + hr + 1] = e_at((row0 + hr) & rmask, (col0 + c) & cmask, nullptr);
    }
    __syncthreads();

    double loss = 0.0;
    const bool row_ok = row0 + r < g.nr;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int c = cg * 8 + k;
        const float* col = &es[(c + 1) * kStLD + r + 1];
        const float e0 = col[0];
        const float el = col[-kStLD], er = col[kStLD], eu = col[-1], ed = col[1];
        const float gval = (sgnf(e0 - el) - sgnf(er - e0) + sgnf(e0 - eu) - sgnf(ed - e0)) * gscale;
        const bool ok = row_ok && (col0 + c < g.nc);
        if (ok) {
            loss += (double)fabsf(er - e0) + (double)fabsf(ed - e0);
            const float2 zz = zc[k];
            const float mag = sqrtf(zz.x * zz.x + zz.y * zz.y);
            float2 o = make_float2(0.f, 0.f);
            if (mag > 0.0f) {
                const float inv = 1.0f / mag;
                o = make_float2(zz.x * inv * gval, zz.y * inv * gval);
            }
            u[page + (size_t)(col0 + c) * g.nr + (row0 + r)] = o;
            if (signs)  // bit 0: sign of er - e0, bit 1: sign of ed - e0 (the copysign decisions above)
                signs[page + (size_t)(col0 + c) * g.nr + (row0 + r)] =
                    (uint8_t)((signbit(er - e0) ? 1 : 0) | (signbit(ed - e0) ? 2 : 0));
        }
    }
    if (loss_acc) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off, 64);
        if ((t & 63) == 0) red[t >> 6] = loss;
        __syncthreads();
        if (t == 0) unsafeAtomicAdd(loss_acc, red[0] + red[1] + red[2] + red[3]);
    }
}

void launch_stencil(const Geom& g, const float2* z, const YStack& y, uint32_t led0, uint32_t b, int pratio,
                    float2* u, double* loss_acc, hipStream_t st, uint8_t* signs) {
    // the reference forms both scales in integer arithmetic then casts
    const uint32_t plane = g.plane();
    const float ratio = (float)(plane * (uint32_t)(pratio * pratio));
    const float gscale = (float)(pratio * pratio);
    dim3 grid((g.nr + kStTR - 1) / kStTR, (g.nc + kStTC - 1) / kStTC, b);
    hipLaunchKernelGGL(stencil_kernel, grid, dim3(256), 0, st, g, z, y, led0, ratio, exact_inverse_or_zero(ratio), gscale,
                       u, loss_acc, signs);
    FPM_LAUNCH_CHECK();
}

// ---------------------------------------------------- back-projection
// cuFPM_v3/cuda/kernels.cuh:67-129 (fused_deconvPIE_shifted) and the pupil
// half of constrained_RMSprop_both (kernels.cuh:384-410).  One thread per
// pupil pixel walks the batch in LED order: the pupil gradient sum stays in a
// register (no atomics, deterministic) and the pupil update needs no second
// kernel, since only this thread reads P at that pixel.  Grouped tiles
// (tab): grid row k is tile k, whose R pages are k b.., whose spectrum, pupil,
// gradient and moment are block k of Shat, P, Gs, mP (GP) and whose LED
// windows and A come from tab[k].
__global__ __launch_bounds__(256) void backproj_kernel(Geom g, const float2* __restrict__ R,
                                                       const float2* __restrict__ Shat, const int2* __restrict__ led,
                                                       uint32_t led0, uint32_t b, float2* __restrict__ P,
                                                       float2* __restrict__ Gs, int mode, float* __restrict__ mP,
                                                       const float* __restrict__ A, RmsArgs rms,
                                                       float2* __restrict__ GP, float invN2,
                                                       const TileIn* __restrict__ tab) {
    const uint32_t plane = g.nr * g.nc;
    const uint32_t pix = blockIdx.x * 256u + threadIdx.x;
    if (pix >= plane) return;
    if (tab) {
        const uint32_t tk = blockIdx.y;
        const size_t NN = (size_t)g.Nr * g.Nc;
        R += (size_t)tk * b * plane;
        Shat += tk * NN;
        P += (size_t)tk * plane;
        if (Gs) Gs += tk * NN;
        if (mP) mP += (size_t)tk * plane;
        if (GP) GP += (size_t)tk * plane;
        led = tab[tk].led;
        A = tab[tk].A;
    }
    const uint32_t i = pix & (g.nr - 1), j = pix >> g.lnr;
    const uint32_t di = (i + (g.nr >> 1)) & (g.nr - 1), dj = (j + (g.nc >> 1)) & (g.nc - 1);
    const size_t rpos = g.nl.off(di, dj);
    const float2 p = P[pix];
    const bool p_zero = p.x == 0.0f && p.y == 0.0f;
    // Where P = 0 the stitch contribution R conj(P) is zero, and where also
    // A = 0 the clamped pupil update below is zero whatever the gradient: the
    // spectra R are then not read (the forward column passes may have skipped
    // those rows).  The gradient-only call (mP == nullptr) computes every pixel.
    const bool skip = mP && p_zero && A[pix] == 0.0f;
    float2 gp = make_float2(0.f, 0.f);
    // unrolled: several LEDs' loads in flight at once (the sums keep LED order)
#pragma unroll 4
    for (uint32_t ell = 0; ell < (skip ? 0u : b); ++ell) {
        const size_t page = (size_t)ell * plane;
        const float2 Rv = R[page + rpos];
        // the window element: where S was cropped from and where the
        // spectrum gradient goes (kernels.cuh:43-47 and :107-113 agree)
        const int2 k = led[led0 + ell];
        const uint32_t srow = (i + (uint32_t)k.y - 1u + (g.Nr >> 1)) & (g.Nr - 1);
        const uint32_t scol = (j + (uint32_t)k.x - 1u + (g.Nc >> 1)) & (g.Nc - 1);
        const size_t spos = (size_t)g.pN(scol) * g.Nr + srow;
        if (mode == kBackprojAtomic && !p_zero) {
            const float2 gw = cmulc(Rv, p);
            float* dst = (float*)(Gs + spos);
            unsafeAtomicAdd(dst, gw.x * invN2);
            unsafeAtomicAdd(dst + 1, gw.y * invN2);
        }
        const float2 gq = cmulc(Rv, Shat[spos]);
        gp.x += gq.x;
        gp.y += gq.y;
    }
    if (mP) {
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

void launch_backproj(const Geom& g, const float2* R, const float2* Shat, const int2* led, uint32_t led0, uint32_t b,
                     float2* P, float2* Gs, int mode, float* mP, const float* A, RmsArgs rms, float2* GP,
                     hipStream_t st, const TileIn* tab, uint32_t ntiles) {
    const float invN2 = 1.0f / (float)(g.Nr * g.Nc);
    if (!tab) ntiles = 1;
    if (ntiles < 1 || ntiles > 65535) throw HipError{hipErrorInvalidValue, "back-projection: 1..65535 tiles"};
    dim3 grid((g.plane() + 255) / 256, ntiles);
    hipLaunchKernelGGL(backproj_kernel, grid, dim3(256), 0, st, g, R, Shat, led, led0, b, P, Gs, mode, mP, A, rms, GP,
                       invN2, tab);
    FPM_LAUNCH_CHECK();
}

// ------------------------------------------------- owner-computes stitch
// Deterministic alternative to the atomics of kernels.cuh:112-116: each
// workgroup owns a 64-row x 16-column tile of the (natural-order) spectrum,
// finds the LED windows that meet the tile, and sums their contributions in
// LED order.  Every element is written (zeros outside all windows), so the
// gradient buffer needs no clearing.
constexpr int kOwTR = 64, kOwTC = 16;
#ifndef FPM_STITCH_PAIRS
#define FPM_STITCH_PAIRS 2  // nonzero: each LED's four column loads issued before their products
#endif

// Grouped tiles (tab): grid z = tile k, its R pages k b.., pupil and
// gradient blocks k, LED windows from tab[k].
__global__ __launch_bounds__(256) void owner_stitch_kernel(Geom g, const float2* __restrict__ R,
                                                           const float2* __restrict__ P, const int2* __restrict__ led,
                                                           uint32_t led0, uint32_t b, float2* __restrict__ Gs,
                                                           float invN2, const TileIn* __restrict__ tab,
                                                           const uint32_t* __restrict__ pcols) {
    __shared__ uint64_t hit[4];
    if (tab) {
        const uint32_t tk = blockIdx.z;
        R += (size_t)tk * b * g.nr * g.nc;
        P += (size_t)tk * g.nr * g.nc;
        Gs += (size_t)tk * g.Nr * g.Nc;
        led = tab[tk].led;
        if (pcols) pcols += (size_t)tk * g.nr;
    }
    const int t = threadIdx.x;
    const int r = t & 63, cg = t >> 6;
    const uint32_t ra = blockIdx.x * kOwTR, ca = blockIdx.y * kOwTC;
    const uint32_t Rm = g.Nr - 1, Cm = g.Nc - 1;
    const uint32_t th = g.Nr < kOwTR ? g.Nr : kOwTR, tw = g.Nc < kOwTC ? g.Nc : kOwTC;
    const uint32_t plane = g.nr * g.nc;
    const uint32_t row = ra + r;
    float2 acc[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[k] = make_float2(0.f, 0.f);

    for (uint32_t c0 = 0; c0 < b; c0 += 256) {
        if (t < 4) hit[t] = 0;
        __syncthreads();
        const uint32_t me = c0 + t;
        if (me < b) {
            const int2 k = led[led0 + me];
            const uint32_t wr = ((uint32_t)k.y - 1u + (g.Nr >> 1)) & Rm;
            const uint32_t wc = ((uint32_t)k.x - 1u + (g.Nc >> 1)) & Cm;
            const bool rows = ((ra - wr) & Rm) < g.nr || ((wr - ra) & Rm) < th;
            const bool cols = ((ca - wc) & Cm) < g.nc || ((wc - ca) & Cm) < tw;
            if (rows && cols) atomicOr((unsigned long long*)&hit[t >> 6], 1ull << (t & 63));
        }
        __syncthreads();
        for (int w = 0; w < 4; ++w) {
            uint64_t m = hit[w];
            while (m) {
                const int bit = __ffsll((unsigned long long)m) - 1;
                m &= m - 1;
                const uint32_t ell = c0 + w * 64 + bit;
                const int2 k = led[led0 + ell];
                const uint32_t wr = ((uint32_t)k.y - 1u + (g.Nr >> 1)) & Rm;
                const uint32_t wc = ((uint32_t)k.x - 1u + (g.Nc >> 1)) & Cm;
                const uint32_t li = (row - wr) & Rm;
                if (li >= g.nr) continue;
                // P = 0 outside pupil row li's support interval [c0, c0 + cw);
                // inside it R is loaded beside P
                uint32_t c0 = 0, cw = g.nc;
                if (pcols) {
                    const uint32_t cc = pcols[li];
                    c0 = cc & 0xffffu;
                    cw = (cc >> 16) - c0;
                }
                const uint32_t di = (li + (g.nr >> 1)) & (g.nr - 1);
#if FPM_STITCH_PAIRS
                // the four columns' R and P loads issued before their products
                float2 Rq[4], Pq[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    Rq[q] = Pq[q] = make_float2(0.f, 0.f);
                    const uint32_t lj = (ca + cg * 4 + q - wc) & Cm;
                    if (lj - c0 < cw) {  // (without intervals R is read wherever P is)
                        Rq[q] = R[(size_t)ell * plane + g.nl.off(di, (lj + (g.nc >> 1)) & (g.nc - 1))];
                        Pq[q] = P[lj * g.nr + li];
                    }
                }
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (Pq[q].x != 0.0f || Pq[q].y != 0.0f) {  // P = 0 contributes +-0
                        const float2 gw = cmulc(Rq[q], Pq[q]);
                        acc[q].x = fmaf(gw.x, invN2, acc[q].x);
                        acc[q].y = fmaf(gw.y, invN2, acc[q].y);
                    }
#else
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint32_t col = ca + cg * 4 + q;
                    const uint32_t lj = (col - wc) & Cm;
                    if (lj - c0 < cw) {
                        const uint32_t dj = (lj + (g.nc >> 1)) & (g.nc - 1);
                        const float2* rp = R + (size_t)ell * plane + g.nl.off(di, dj);
                        float2 Rv = make_float2(0.f, 0.f);
                        if (pcols) Rv = *rp;
                        const float2 pp = P[lj * g.nr + li];
                        if (pp.x != 0.0f || pp.y != 0.0f) {  // P = 0 contributes +-0
                            if (!pcols) Rv = *rp;
                            const float2 gw = cmulc(Rv, pp);
                            acc[q].x = fmaf(gw.x, invN2, acc[q].x);
                            acc[q].y = fmaf(gw.y, invN2, acc[q].y);
                        }
                    }
                }
#endif
            }
        }
        __syncthreads();
    }
    if (row < g.Nr) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t col = ca + cg * 4 + q;
            if (col < g.Nc) Gs[(size_t)g.pN(col) * g.Nr + row] = acc[q];
        }
    }
}

void launch_owner_stitch(const Geom& g, const float2* R, const float2* P, const int2* led, uint32_t led0, uint32_t b,
                         float2* Gs, hipStream_t st, const TileIn* tab, uint32_t ntiles, const uint32_t* pcols) {
    const float invN2 = 1.0f / (float)(g.Nr * g.Nc);
    if (!tab) ntiles = 1;
    if (ntiles < 1 || ntiles > 65535) throw HipError{hipErrorInvalidValue, "owner stitch: 1..65535 tiles"};
    dim3 grid((g.Nr + kOwTR - 1) / kOwTR, (g.Nc + kOwTC - 1) / kOwTC, ntiles);
    hipLaunchKernelGGL(owner_stitch_kernel, grid, dim3(256), 0, st, g, R, P, led, led0, b, Gs, invN2, tab, pcols);
    FPM_LAUNCH_CHECK();
}

// ------------------------------------------------------ window row masks
// mask[t] = 1 iff rows [16t, 16t+16) of the N-side spectrum meet the window
// of any LED led0..led0+b-1 at a pupil row i with sup[i] != 0 (window row
// (kt - 1 + Nr/2 + i) mod Nr, i < nr; sup == nullptr: every pupil row).
// Outside those rows the spectrum gradient of the batch is exactly zero (the
// windows' P rows there are all 0) and the spectrum is never read (the crop
// and the pupil kernel read it only where P != 0 or A != 0), so the column
// passes skip them.
// One workgroup = kRmEntries mask entries x 32 LED lanes: lane j tests LEDs
// j, j + 32 (b <= 64 in a solve; more loop) against the 16 rows of its entry,
// with the pupil support staged in LDS, and a ballot ORs the lanes.  Grouped
// tiles (tab): grid row k is tile k, its LED windows from tab[k], its support
// at sup + k nr, its mask at mask + k Nr/16.
constexpr int kRmEntries = 8;
constexpr int kRmMaxSup = 16384;  // pupil rows (the largest side)
__global__ __launch_bounds__(256) void row_mask_kernel(const int2* __restrict__ led, uint32_t led0, uint32_t b,
                                                       uint32_t Nr, uint32_t nr, const uint8_t* __restrict__ sup,
                                                       uint8_t* __restrict__ mask, const TileIn* __restrict__ tab,
                                                       int axis) {
    __shared__ uint8_t ssup[kRmMaxSup];
    if (tab) {
        const uint32_t tk = blockIdx.y;
        led = tab[tk].led;
        if (sup) sup += (size_t)tk * nr;
        mask += (size_t)tk * (Nr / kMaskRows);
    }
    if (sup)
        for (uint32_t i = threadIdx.x; i < nr; i += 256) ssup[i] = sup[i];
    __syncthreads();
    const uint32_t e = blockIdx.x * kRmEntries + threadIdx.x / 32;
    const uint32_t lane = threadIdx.x % 32;
    const uint32_t Rm = Nr - 1, ra = e * kMaskRows;
    bool hit = false;
    if (e < Nr / kMaskRows) {
        for (uint32_t l = lane; l < b; l += 32) {
            const int2 k = led[led0 + l];
            const uint32_t wr = ((uint32_t)(axis ? k.x : k.y) - 1u + (Nr >> 1)) & Rm;
            for (uint32_t r = 0; r < kMaskRows; ++r) {
                const uint32_t i = (ra + r - wr) & Rm;
                hit |= i < nr && (!sup || ssup[i]);
            }
        }
    }
    const uint64_t bal = __ballot(hit);
    if (e < Nr / kMaskRows && lane == 0) mask[e] = ((bal >> (threadIdx.x & 32u)) & 0xffffffffull) != 0 ? 1 : 0;
}

void launch_row_mask(const int2* led, uint32_t led0, uint32_t b, uint32_t Nr, uint32_t nr, const uint8_t* sup,
                     uint8_t* mask, hipStream_t st, const TileIn* tab, uint32_t ntiles, int axis) {
    const uint32_t entries = Nr / kMaskRows;
    if (!tab) ntiles = 1;
    if (ntiles < 1 || ntiles > 65535) throw HipError{hipErrorInvalidValue, "row mask: 1..65535 tiles"};
    if (nr > (uint32_t)kRmMaxSup) throw HipError{hipErrorInvalidValue, "row mask: pupil side above 16384"};
    hipLaunchKernelGGL(row_mask_kernel, dim3((entries + kRmEntries - 1) / kRmEntries, ntiles), dim3(256), 0, st, led,
                       led0, b, Nr, nr, sup, mask, tab, axis);
    FPM_LAUNCH_CHECK();
}

// csup[j] = 1 iff pupil column j lies in the support interval of some pupil
// row (cols from launch_pupil_row_mask): a superset of the columns holding a
// P != 0 or an A != 0, at every step of a solve.
__global__ __launch_bounds__(256) void pupil_col_support_kernel(const uint32_t* __restrict__ cols, uint32_t nr,
                                                                uint32_t nc, uint8_t* __restrict__ csup) {
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j >= nc) return;
    uint8_t hit = 0;
    for (uint32_t i = 0; i < nr; ++i) {
        const uint32_t cc = cols[i], c0 = cc & 0xffffu;
        hit |= (j - c0 < (cc >> 16) - c0) ? 1 : 0;
    }
    csup[j] = hit;
}

void launch_pupil_col_support(const uint32_t* cols, uint32_t nr, uint32_t nc, uint8_t* csup, hipStream_t st) {
    hipLaunchKernelGGL(pupil_col_support_kernel, dim3((nc + 255) / 256), dim3(256), 0, st, cols, nr, nc, csup);
    FPM_LAUNCH_CHECK();
}

// Row masks from the pupil support.  sup[i] = 1 iff pupil row i holds a
// P != 0 or an A != 0.  The per-LED spectra are ifftshifted (spectrum row di
// holds pupil row i = (di + nr/2) mod nr): mask[t] = 1 iff spectrum rows
// [16t, 16t+16) hold a pupil row with sup = 1.  Outside them S P is zero (the
// inverse column passes skip those rows) and R is never read (P = 0
// contributes nothing to the stitch, and where A = 0 the clamped pupil update
// is 0 whatever the gradient, so the forward column passes skip them too).
// Since every update leaves P = 0 wherever A = 0, the masks of the initial
// pupil hold for every later step.
__global__ __launch_bounds__(1024) void pupil_row_mask_kernel(const float2* __restrict__ P,
                                                              const float* __restrict__ A, uint32_t nr, uint32_t nc,
                                                              uint8_t* __restrict__ mask, uint8_t* __restrict__ sup,
                                                              uint32_t* __restrict__ cols) {
    __shared__ int any[kMaskRows];
    __shared__ uint32_t lo[kMaskRows], hi[kMaskRows];
    const uint32_t t = threadIdx.x, tile = blockIdx.x;
    if (t < kMaskRows) {
        any[t] = 0;
        lo[t] = 0xffffffffu;
        hi[t] = 0;
    }
    __syncthreads();
    // 1024 is a multiple of 16: every thread stays on spectrum row t mod 16
    const uint32_t di = tile * kMaskRows + t % kMaskRows;
    const uint32_t i = (di + (nr >> 1)) & (nr - 1);
    int h = 0;
    uint32_t jlo = 0xffffffffu, jhi = 0;
    for (uint32_t j = t / kMaskRows; j < nc; j += 1024 / kMaskRows) {
        const float2 p = P[(size_t)j * nr + i];
        if ((p.x != 0.0f) | (p.y != 0.0f) | (A[(size_t)j * nr + i] != 0.0f)) {
            h = 1;
            jlo = min(jlo, j);
            jhi = j + 1;
        }
    }
    if (h) {
        atomicOr(&any[t % kMaskRows], 1);
        atomicMin(&lo[t % kMaskRows], jlo);
        atomicMax(&hi[t % kMaskRows], jhi);
    }
    __syncthreads();
    if (t < kMaskRows) {
        const uint32_t row = (tile * kMaskRows + t + (nr >> 1)) & (nr - 1);
        sup[row] = (uint8_t)any[t];
        if (cols) cols[row] = any[t] ? lo[t] | hi[t] << 16 : 0u;
    }
    if (t == 0) {
        int m = 0;
        for (int r = 0; r < kMaskRows; ++r) m |= any[r];
        mask[tile] = (uint8_t)m;
    }
}

void launch_pupil_row_mask(const float2* P, const float* A, uint32_t nr, uint32_t nc, uint8_t* mask, uint8_t* sup,
                           hipStream_t st, uint32_t* cols) {
    if (cols && nc > 65535u) throw HipError{hipErrorInvalidValue, "pupil column intervals: pupil side above 65535"};
    hipLaunchKernelGGL(pupil_row_mask_kernel, dim3(nr / kMaskRows), dim3(1024), 0, st, P, A, nr, nc, mask, sup, cols);
    FPM_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- helpers
// a fill kernel rather than hipMemsetAsync: the call may sit in a captured graph
__global__ __launch_bounds__(256) void zero_kernel(float4* __restrict__ p, size_t n4, float2* __restrict__ tail) {
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n4; k += (size_t)gridDim.x * 256)
        p[k] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (tail && blockIdx.x == 0 && threadIdx.x == 0) *tail = make_float2(0.f, 0.f);
}

void launch_zero(float2* p, size_t count, hipStream_t st) {
    // count complex = count/2 float4 (+ one float2 when count is odd)
    const size_t n4 = count / 2;
    const size_t blocks = std::max<size_t>(1, std::min<size_t>((n4 + 255) / 256, 1u << 20));
    if (count) hipLaunchKernelGGL(zero_kernel, dim3((unsigned)blocks), dim3(256), 0, st, (float4*)p, n4,
                                  (count & 1) ? p + count - 1 : nullptr);
    FPM_LAUNCH_CHECK();
}

__global__ __launch_bounds__(256) void abs_scaled_kernel(const float2* __restrict__ z, size_t n, float scale,
                                                         int squared, float* __restrict__ out) {
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256) {
        const float2 v = z[k];
        const float a = sqrtf(v.x * v.x + v.y * v.y) / scale;
        out[k] = squared ? a * a : a;
    }
}

void launch_abs_scaled(const float2* z, size_t count, float scale, int squared, float* out, hipStream_t st) {
    const size_t blocks = std::min<size_t>((count + 255) / 256, 1u << 20);
    hipLaunchKernelGGL(abs_scaled_kernel, dim3((unsigned)blocks), dim3(256), 0, st, z, count, scale, squared, out);
    FPM_LAUNCH_CHECK();
}

// min/max of a uint16 stack: 16-byte loads (8 values) over the aligned body,
// single values for the unaligned head and the tail, a workgroup reduction
// in LDS and one atomic pair per workgroup
__device__ __forceinline__ void mm_u32pair(uint32_t w, uint32_t& lo, uint32_t& hi) {
    const uint32_t a = w & 0xffffu, b = w >> 16;
    lo = min(lo, min(a, b));
    hi = max(hi, max(a, b));
}

__global__ __launch_bounds__(256) void u16_minmax_kernel(const uint16_t* __restrict__ raw, size_t n, size_t head,
                                                         size_t nvec, uint32_t* __restrict__ mm) {
    __shared__ uint32_t slo[4], shi[4];
    uint32_t lo = 0xffffu, hi = 0u;
    const size_t tid = (size_t)blockIdx.x * 256 + threadIdx.x, stride = (size_t)gridDim.x * 256;
    const uint4* __restrict__ body = (const uint4*)(raw + head);
    for (size_t k = tid; k < nvec; k += stride) {
        const uint4 v = body[k];
        mm_u32pair(v.x, lo, hi);
        mm_u32pair(v.y, lo, hi);
        mm_u32pair(v.z, lo, hi);
        mm_u32pair(v.w, lo, hi);
    }
    // head (before the first 16-byte boundary) and tail (after the last full vector)
    for (size_t k = tid; k < head; k += stride) {
        const uint32_t v = raw[k];
        lo = min(lo, v);
        hi = max(hi, v);
    }
    for (size_t k = head + nvec * 8 + tid; k < n; k += stride) {
        const uint32_t v = raw[k];
        lo = min(lo, v);
        hi = max(hi, v);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, (uint32_t)__shfl_down(lo, off, 64));
        hi = max(hi, (uint32_t)__shfl_down(hi, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        slo[threadIdx.x >> 6] = lo;
        shi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) {
            lo = min(lo, slo[w]);
            hi = max(hi, shi[w]);
        }
        atomicMin(&mm[0], lo);
        atomicMax(&mm[1], hi);
    }
}

void launch_u16_minmax(const uint16_t* raw, size_t count, uint32_t* minmax, hipStream_t st) {
    if ((uintptr_t)raw & 1u) throw HipError{hipErrorInvalidValue, "uint16 image stack is not 2-byte aligned"};
    const size_t head = std::min<size_t>(count, ((16u - ((uintptr_t)raw & 15u)) & 15u) / 2u);
    const size_t nvec = (count - head) / 8;
    const size_t blocks = std::max<size_t>(1, std::min<size_t>((nvec + 255) / 256, 2048));
    hipLaunchKernelGGL(u16_minmax_kernel, dim3((unsigned)blocks), dim3(256), 0, st, raw, count, head, nvec, minmax);
    FPM_LAUNCH_CHECK();
}

__global__ __launch_bounds__(256) void u16_to_amp_kernel(const uint16_t* __restrict__ raw, size_t n, uint32_t mn,
                                                         float range, float rcp_range, float* __restrict__ out) {
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256)
        out[k] = sqrtf(div_u16((float)((uint32_t)raw[k] - mn), range, rcp_range));
}

void launch_u16_to_amp(const uint16_t* raw, size_t count, uint32_t mn, float range, float* out, hipStream_t st) {
    const size_t blocks = std::min<size_t>((count + 255) / 256, 1u << 20);
    hipLaunchKernelGGL(u16_to_amp_kernel, dim3((unsigned)blocks), dim3(256), 0, st, raw, count, mn, range, 1.0f / range,
                       out);
    FPM_LAUNCH_CHECK();
}

// ------------------------------------------------------------ solve state
__global__ void solve_state_init_kernel(SolveState s, float lr0) {
    if (threadIdx.x == 0) {
        *s.lr = lr0;
        *s.epoch = 0u;
        if (s.loss_cur)
            for (uint32_t k = 0; k < s.ntiles; ++k) s.loss_cur[k] = 0.0;
    }
}

__global__ void epoch_end_kernel(SolveState s) {
    if (threadIdx.x == 0) {
        const uint32_t e = *s.epoch;
        if (s.loss_cur) {
            for (uint32_t k = 0; k < s.ntiles; ++k) {
                if (s.loss_hist) s.loss_hist[(size_t)k * s.hist_stride + e] = s.loss_cur[k];
                s.loss_cur[k] = 0.0;
            }
        }
        // float lr times double 0.5, as the reference's host loop does
        if (e >= s.halve_from) *s.lr = (float)((double)*s.lr * 0.5);
        *s.epoch = e + 1u;
    }
}

void launch_solve_state_init(const SolveState& s, float lr0, hipStream_t st) {
    hipLaunchKernelGGL(solve_state_init_kernel, dim3(1), dim3(64), 0, st, s, lr0);
    FPM_LAUNCH_CHECK();
}

void launch_epoch_end(const SolveState& s, hipStream_t st) {
    hipLaunchKernelGGL(epoch_end_kernel, dim3(1), dim3(64), 0, st, s);
    FPM_LAUNCH_CHECK();
}

}  // namespace fpm
