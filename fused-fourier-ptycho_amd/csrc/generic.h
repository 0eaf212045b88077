// generic.h -- the solve's steps for even sides that are not powers of two
// (generic.hip): mixed-radix transforms and unfused kernels in natural order.
#pragma once
#include "common.h"
#include "fpm_kernels.h"

namespace fpm {

// largest side of the generic path (its FFT holds a whole line, up to 16384
// points = 128 KiB, in LDS)
constexpr int kGenericMaxSide = 16384;

// even v in [16, kGenericMaxSide] (powers of two included)
bool generic_side(int64_t v);

// unnormalised 2D C2C transforms of `batch` column-major rows x cols pages,
// natural order in and out (in == out allowed); rows, cols generic sides
void fft2_generic(const float2* in, float2* out, uint32_t rows, uint32_t cols, uint32_t batch, bool inverse,
                  hipStream_t st);
// the twiddle tables of the geometry's four lengths, made before any capture
void generic_prepare(const Geom& g);
// Launch shapes of the solve's transform passes (T tiles, batch B): each
// pass not yet tuned on this device is timed with candidate shapes (points
// per workgroup x points per thread) on the given buffers -- O (T objects),
// S (T spectra), P, led or tab, pages (T B per-LED pages), read only; the
// outputs go to a scratch buffer -- and the fastest is kept for every later
// launch of that (length, line kind, source).  Results do not depend on
// the shape.  Synchronises st; does nothing while st is capturing.
void generic_tune(const Geom& g, uint32_t T, uint32_t B, const float2* O, const float2* S, const float2* P,
                  const int2* led, const TileIn* tab, const float2* pages, hipStream_t st);

// crop x pupil of LEDs led0.. into b ifftshifted natural pages
// tab (optional, ntiles entries, ntiles <= 65535): grouped tiles, tile k's
// arrays at block k of each argument (pages: b per tile), its LED windows,
// images, mask and loss slot from tab[k]; led / y / A / loss_acc unused then
void generic_crop(const float2* S, const float2* P, const int2* led, uint32_t led0, uint32_t b, const Geom& g,
                  float2* out, hipStream_t st, const TileIn* tab = nullptr, uint32_t ntiles = 1);
// residual (ebuf: b n_r n_c floats) and loss stencil: u from z; loss_acc may be null;
// signs (optional, not with tab): the stencil's decisions, one byte per pixel
// (bit 0: sign bit of e(i, j+1) - e(i, j), bit 1: of e(i+1, j) - e(i, j))
void generic_stencil(const float2* z, const YStack& y, uint32_t led0, uint32_t b, int pratio, const Geom& g,
                     float* ebuf, float2* u, double* loss_acc, hipStream_t st, const TileIn* tab = nullptr,
                     uint32_t ntiles = 1, uint8_t* signs = nullptr);
// The per-LED pages of a minibatch step, in three stages: lat = IDFT_n(crop x
// pupil), rec = the loss stencil's u, then rec = DFT_n(rec).  fused
// (generic_fused_pages): the crop is the source of the first row pass, and
// the inverse column pass, residual, stencil and forward column pass run as
// one kernel on column sets (lat holds the pages between the row pass and
// the columns; ebuf unused); else generic_crop, fft2_generic,
// generic_stencil, fft2_generic.  The same arithmetic per element either
// way, with the row and column passes in the other order.
bool generic_fused_pages(const Geom& g, bool signs);
void generic_inverse_pages(const float2* S, const float2* P, const int2* led, uint32_t led0, uint32_t b,
                           const Geom& g, float2* lat, hipStream_t st, const TileIn* tab, uint32_t ntiles,
                           bool fused);
void generic_stencil_pages(const float2* z, const YStack& y, uint32_t led0, uint32_t b, int pratio, const Geom& g,
                           float* ebuf, float2* u, double* loss_acc, hipStream_t st, const TileIn* tab,
                           uint32_t ntiles, uint8_t* signs, bool fused);
void generic_forward_pages(float2* rec, uint32_t pages, const Geom& g, hipStream_t st, bool fused);
// owner-computes stitch of the b back-projections into G_S (every element
// written); pcols (optional): pupil-support column intervals (generic_pupil_cols)
void generic_stitch(const float2* R, const float2* P, const int2* led, uint32_t led0, uint32_t b, const Geom& g,
                    float2* Gs, hipStream_t st, const TileIn* tab = nullptr, uint32_t ntiles = 1,
                    const uint32_t* pcols = nullptr);
// cols (nr words per tile) = the smallest column interval of each pupil row
// holding every P != 0 and A != 0 (A per tile from tab when grouped)
void generic_pupil_cols(const float2* P, const float* A, const Geom& g, uint32_t* cols, hipStream_t st,
                        const TileIn* tab = nullptr, uint32_t ntiles = 1);
// pupil gradient; with mP and A the RMSprop + clamp update of P, else GP = gradient
void generic_pupil(const float2* R, const float2* S, const int2* led, uint32_t led0, uint32_t b, const Geom& g,
                   float2* P, float* mP, const float* A, RmsArgs rms, float2* GP, hipStream_t st,
                   const TileIn* tab = nullptr, uint32_t ntiles = 1);
// RMSprop of O from its gradient G (n elements)
void generic_update(const float2* G, float2* O, float* mO, size_t n, RmsArgs rms, hipStream_t st);
// the strided-index pass of a 2D transform alone (in place)
void generic_strided_pass(float2* data, uint32_t rows, uint32_t cols, uint32_t batch, bool inverse, hipStream_t st);
// N_rows = 3 x 2^k: the object update fused with the contiguous passes --
// G (ntiles gradient spectra after their strided inverse pass) -> G_O along
// the contiguous index, RMSprop of O, and with forward the contiguous pass
// of DFT_N(O) written back into G (its strided pass then completes the next
// step's spectrum)
bool generic_update_fusable(const Geom& g);
void generic_update_fused(float2* G, float2* O, float* mO, const Geom& g, uint32_t ntiles, RmsArgs rms, bool forward,
                          hipStream_t st);

}  // namespace fpm
