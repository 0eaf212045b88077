// fft_kernels.h -- host launchers of the HIP FFT passes.
#pragma once
#include "common.h"

namespace fpm {

// Crop of LED windows from the column-permuted N-side spectrum times the
// pupil, formed inside the first inverse column pass of the batched transform
// (see fft_cols_inverse).  Page l of the batch is LED led0 + l.
// tab (optional): grouped tiles, ppt pages per tile; page k ppt + l is LED
// led0 + l of tile k, whose spectrum and pupil are block k of Shat and P and
// whose LED windows come from tab[k] (led unused).
struct CropArgs {
    const float2* Shat;
    const float2* P;
    const int2* led;
    uint32_t led0;
    uint32_t nr, nc, Nr, Nc;
    ColPerm pN;
    const TileIn* tab = nullptr;
    uint32_t ppt = 0;
    // optional: pupil row i holds P != 0 only in columns [c0, c1), pcols[i] =
    // c0 | c1 << 16 (launch_pupil_row_mask); grouped tiles: nr words per tile
    const uint32_t* pcols = nullptr;
};

// Owner-computes stitch formed inside the first inverse column pass of the
// object-spectrum gradient: element (row, natural col) = sum over the LEDs
// led0..led0+b-1 (b <= 64) whose window covers it, in LED order, of
// R_l conj(P) / (Nr Nc), R = the column-permuted per-LED spectra.
// tab (optional): grouped tiles, page k of the transform is tile k's
// gradient, summed from R pages k b.., pupil block k of P and the LED
// windows of tab[k] (led unused).
struct StitchArgs {
    const float2* R;
    const float2* P;
    const int2* led;
    uint32_t led0, b;
    uint32_t nr, nc, Nr, Nc;
    PageLayout nl;  // layout of the R pages
    float invN2;
    const TileIn* tab = nullptr;
    // optional: pupil row i holds P != 0 only in columns [c0, c1), pcols[i] =
    // c0 | c1 << 16 (launch_pupil_row_mask); grouped tiles: nr words per tile
    const uint32_t* pcols = nullptr;
};

// one LEN-point transform per contiguous line (LEN = len, nlines lines)
// eng: ENG_F32 or ENG_SPLIT16 (every launcher below)
void fft_contig(const float2* in, float2* out, uint32_t len, uint32_t nlines, bool inverse, const float2* twid,
                hipStream_t st, int eng = ENG_F32);

// transforms along the columns index of rows x cols pages; see fft_kernels.hip.
// rowmask (optional, one byte per 16 rows): row tiles whose bytes are all 0
// are skipped -- their output is left as it was (forward: rows nobody reads;
// inverse: rows whose input and so output are zero, masked again by the reader).
// lay (optional): the pages' layout (default column-major); a tiled layout
// needs p = identity (one pass, cols <= 2048).
// mask_ppt > 0: one row mask (rows/16 bytes) per mask_ppt pages, back to back
// (grouped tiles); 0: the one mask serves every page.
// colmask (optional, one page): the last pass stores only the natural
// columns whose 16-column group has colmask != 0 (Nc / 16 bytes)
// pcols (optional, one pass): pupil-support column intervals c0 | c1 << 16 of
// the rows (launch_pupil_row_mask's, indexed by the pupil row of the
// ifftshifted page and tested on the pupil column); each row stores only the
// points inside its interval (with mask_ppt: rows intervals per mask_ppt pages)
void fft_cols_forward(float2* data, uint32_t rows, uint32_t cols, uint32_t batch, const ColPerm& p, const float2* twid,
                      hipStream_t st, float2* nat_out, const uint8_t* rowmask = nullptr,
                      const PageLayout* lay = nullptr, int eng = ENG_F32, uint32_t mask_ppt = 0,
                      const uint8_t* colmask = nullptr, const uint32_t* pcols = nullptr);
// With crop, the input is not read: the first pass gathers ifftshift(S P)
// of each page's LED window (natural frequency columns) instead.
void fft_cols_inverse(float2* data, uint32_t rows, uint32_t cols, uint32_t batch, const ColPerm& p, const float2* twid,
                      hipStream_t st, const float2* nat_in, const CropArgs* crop = nullptr,
                      const StitchArgs* stitch = nullptr, const uint8_t* rowmask = nullptr,
                      const PageLayout* lay = nullptr, int eng = ENG_F32, uint32_t mask_ppt = 0);

// full 2D transforms, in place; the frequency side is column-permuted by p
void fft2_forward(float2* data, uint32_t rows, uint32_t cols, uint32_t batch, const ColPerm& p, const float2* twid,
                  hipStream_t st, int eng = ENG_F32);
void fft2_inverse(float2* data, uint32_t rows, uint32_t cols, uint32_t batch, const ColPerm& p, const float2* twid,
                  hipStream_t st, int eng = ENG_F32);

}  // namespace fpm
