// fft_kernels.h -- host launchers of the HIP FFT passes.
#pragma once
#include "common.h"

namespace fpm {

// one LEN-point transform per contiguous line (LEN = len, nlines lines)
void fft_contig(const float2* in, float2* out, uint32_t len, uint32_t nlines, bool inverse, const float2* twid,
                hipStream_t st);

// transforms along the columns index of rows x cols pages; see fft_kernels.hip
void fft_cols_forward(float2* data, uint32_t rows, uint32_t cols, uint32_t batch, const ColPerm& p, const float2* twid,
                      hipStream_t st, float2* nat_out);
void fft_cols_inverse(float2* data, uint32_t rows, uint32_t cols, uint32_t batch, const ColPerm& p, const float2* twid,
                      hipStream_t st, const float2* nat_in);

// full 2D transforms, in place; the frequency side is column-permuted by p
void fft2_forward(float2* data, uint32_t rows, uint32_t cols, uint32_t batch, const ColPerm& p, const float2* twid,
                  hipStream_t st);
void fft2_inverse(float2* data, uint32_t rows, uint32_t cols, uint32_t batch, const ColPerm& p, const float2* twid,
                  hipStream_t st);

}  // namespace fpm
