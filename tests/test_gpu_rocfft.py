"""The library's 2D DFT against rocFFT (torch.fft on ROCm) on device buffers.

The reference delegates every transform of the solve to cuFFT plans
(cuFPM_v3/cuFPM_pure_v3.cu:118-126, executed at :128-187); rocFFT is the
vendor library of the same contract on this GPU, so the two FP32 results must
agree to within the rounding of either.  A column-major rows x cols page is a
contiguous (cols, rows) torch tensor, and the 2D DFT commutes with the
transpose, so torch.fft.fft2 over the last two dims transforms the same memory.
"""
import pytest
import torch

import fpm_amd

pytestmark = pytest.mark.gpu


def _pages(rows, cols, batch, seed):
    g = torch.Generator(device="cuda:0").manual_seed(seed)
    re = torch.randn((batch, cols, rows), generator=g, device="cuda:0")
    im = torch.randn((batch, cols, rows), generator=g, device="cuda:0")
    return torch.complex(re, im)


def _library_fft2(x, inverse=False):
    out = torch.empty_like(x)
    batch, cols, rows = x.shape
    fpm_amd._check(fpm_amd.lib().fpm_fft2_engine(x.data_ptr(), out.data_ptr(), rows, cols, batch,
                                                 1 if inverse else -1, fpm_amd.FFT_F32, 1,
                                                 torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return out


def _rel(a, b):
    d = (a - b).abs().double().pow(2).sum()
    return (d / b.abs().double().pow(2).sum()).sqrt().item()


# FP32 transforms of N points carry ~ sqrt(log2 N) ulps of relative error each
TOL = 2e-6


@pytest.mark.parametrize("rows,cols,batch", [
    (2048, 2048, 26),     # the per-LED transforms of one minibatch
    (16384, 16384, 1),    # the object spectrum
    (512, 4096, 2),       # non-square, four-step columns
    (768, 768, 3),        # generic sides: 2^8 3
    (1600, 1600, 1),      # 2^6 5^2
    (12288, 12288, 1),    # 2^12 3
    (1400, 200, 2),       # 2^3 5^2 7 x 2^3 5^2
])
def test_fft2_matches_rocfft(rows, cols, batch):
    x = _pages(rows, cols, batch, seed=rows + cols + batch)
    fwd = _library_fft2(x)
    ref = torch.fft.fft2(x)
    assert _rel(fwd, ref) < TOL
    inv = _library_fft2(ref, inverse=True)
    ref_inv = torch.fft.ifft2(ref, norm="forward")  # unnormalised inverse
    assert _rel(inv, ref_inv) < TOL
