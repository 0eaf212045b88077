"""Seeded synthetic FPM datasets (the caller side of the solve).

The reference ships no data and no generator for this setting (its datasets are
external, README.md:17,23; benchmark/gen_img_batch.m writes amplitudes of a
ring-LED model).  This generator produces the workload the solve is measured
on:

* optics: wavelength 0.532 um, NA 0.1, effective pixel 1.625 um, 19 x 19 LEDs
  at 4 mm pitch, 90 mm high, centre-out spiral order; LED windows from
  :func:`fpm_amd.led_offsets` (FP64, init_environment_rgb.m:33-134);
* pupil mask A: the NA disk on the low-res frequency grid
  (init_environment_rgb.m:33-40); true pupil A exp(i 0.5 sqrt3 (2 rho^2 - 1))
  (0.5 rad of defocus), or A exp(i Z) for given Zernike coefficients
  (:func:`zernike_pupil`);
* object (N x N): amplitude = uniform noise low-passed by a disk at
  NA 0.1 + max illumination NA, scaled to [0.3, 1]; phase = sum of 20
  Gaussians (centres uniform, sigma in [N/64, N/16], heights in [-pi/2, pi/2]);
* images: I = amp^2 of the solver's image model (fpm_forward_intensity),
  scaled to a brightfield peak of 40000, Poisson noise, + 100, clipped to
  uint16;
* init: O0 = bilinear upsampling of the prepared centre amplitude, P0 = A.

Two backends: ``numpy`` (any size the host handles; the image model is
evaluated with numpy's FFT) and ``device`` (torch tensors on the GPU, the
image model and transforms through libfpm_amd) for the large workloads.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import led_offsets

WAVELENGTH_UM = 0.532
NA = 0.1
PIXEL_UM = 1.625
LED_PITCH_UM = 4000.0
LED_HEIGHT_UM = 90000.0
LED_GRID = 19
PEAK_COUNTS = 40000.0
OFFSET_COUNTS = 100.0


@dataclass
class Dataset:
    n: int
    pratio: int
    led: np.ndarray          # int32 2 x L {kl; kt}
    A: np.ndarray            # n x n float32 pupil mask
    P_true: np.ndarray       # n x n complex64
    O_true: Optional[np.ndarray]  # N x N complex64 (numpy backend only)
    raw: object              # n x n x L uint16 (numpy array or device tensor view)
    amp: object              # n x n x L float32 prepared amplitudes
    O0: object               # N x N complex64 initial object
    P0: np.ndarray           # n x n complex64 initial pupil
    na_illum: np.ndarray


def freq_grid(n: int, pixel_um: float = PIXEL_UM):
    """Centred angular-frequency grid of an n x n image (init_environment_rgb.m:33-35)."""
    f = np.arange(-(n // 2), n // 2, dtype=np.float64) / (n * pixel_um) * (2 * np.pi)
    Fx, Fy = np.meshgrid(f, f)
    return Fx, Fy


def pupil(n: int, defocus_rad: float = 0.5):
    """(A, P_true): NA disk (init_environment_rgb.m:39-40) and a defocused pupil."""
    Fx, Fy = freq_grid(n)
    kmax = NA * 2 * np.pi / WAVELENGTH_UM
    r2 = Fx * Fx + Fy * Fy
    A = (r2 <= kmax * kmax).astype(np.float32)
    rho2 = r2 / (kmax * kmax)
    P = A * np.exp(1j * defocus_rad * np.sqrt(3.0) * (2 * rho2 - 1))
    # the solve's pupil is unshifted-centred like the reference's (centre = n/2+1)
    return np.asfortranarray(A), np.asfortranarray(P.astype(np.complex64))


ZERNIKE_TERMS = ("defocus", "astigmatism 0", "astigmatism 45", "coma x", "coma y")


def zernike_basis(n: int) -> np.ndarray:
    """(5, n, n) float64 phase maps (Noll-normalised over the NA disk, rho =
    |k| / k_max; the pupil mask A cuts them off), in :data:`ZERNIKE_TERMS`
    order: sqrt3 (2 rho^2 - 1),
    sqrt6 rho^2 cos 2t, sqrt6 rho^2 sin 2t, sqrt8 (3 rho^3 - 2 rho) cos t,
    sqrt8 (3 rho^3 - 2 rho) sin t, on the grid of :func:`freq_grid`."""
    Fx, Fy = freq_grid(n)
    kmax = NA * 2 * np.pi / WAVELENGTH_UM
    rho = np.sqrt(Fx * Fx + Fy * Fy) / kmax
    th = np.arctan2(Fy, Fx)
    return np.stack([np.sqrt(3) * (2 * rho ** 2 - 1), np.sqrt(6) * rho ** 2 * np.cos(2 * th),
                  np.sqrt(6) * rho ** 2 * np.sin(2 * th), np.sqrt(8) * (3 * rho ** 3 - 2 * rho) * np.cos(th),
                  np.sqrt(8) * (3 * rho ** 3 - 2 * rho) * np.sin(th)])


def zernike_pupil(n: int, coeffs, A: np.ndarray) -> np.ndarray:
    """A exp(i sum_k c_k Z_k) (n x n complex64, column-major like :func:`pupil`)."""
    Z = np.tensordot(np.asarray(coeffs, dtype=np.float64), zernike_basis(n), axes=1)
    return np.asfortranarray((A * np.exp(1j * Z)).astype(np.complex64))


def bilinear_upsample(img: np.ndarray, r: int) -> np.ndarray:
    """Bilinear upsampling by an integer factor, pixel-centre aligned
    (x_in = (x_out + 0.5)/r - 0.5, clamped at the borders)."""
    n0, n1 = img.shape

    def axis(n):
        x = (np.arange(n * r) + 0.5) / r - 0.5
        x = np.clip(x, 0, n - 1)
        i0 = np.floor(x).astype(np.int64)
        i1 = np.minimum(i0 + 1, n - 1)
        w = x - i0
        return i0, i1, w

    r0, r1, wr = axis(n0)
    c0, c1, wc = axis(n1)
    a = img[np.ix_(r0, c0)] * (1 - wr)[:, None] * (1 - wc)[None, :]
    a += img[np.ix_(r1, c0)] * wr[:, None] * (1 - wc)[None, :]
    a += img[np.ix_(r0, c1)] * (1 - wr)[:, None] * wc[None, :]
    a += img[np.ix_(r1, c1)] * wr[:, None] * wc[None, :]
    return a


def object_numpy(N: int, pratio: int, na_illum_max: float, seed: int) -> np.ndarray:
    """Seeded synthetic object (N x N complex128, Fortran order)."""
    rng = np.random.Generator(np.random.Philox(seed))
    noise = rng.random((N, N))
    # band limit: disk of radius (NA + NA_illum_max) k0 on the high-res grid,
    # whose pixel is PIXEL_UM / pratio
    k0 = 2 * np.pi / WAVELENGTH_UM
    f = np.fft.fftfreq(N, d=PIXEL_UM / pratio) * 2 * np.pi
    FX, FY = np.meshgrid(f, f)
    cut = (NA + na_illum_max) * k0
    spec = np.fft.fft2(noise) * ((FX * FX + FY * FY) <= cut * cut)
    amp = np.real(np.fft.ifft2(spec))
    amp = 0.3 + 0.7 * (amp - amp.min()) / (amp.max() - amp.min())
    yy, xx = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    phase = np.zeros((N, N))
    for _ in range(20):
        cy, cx = rng.random(2) * N
        sig = N / 64 + rng.random() * (N / 16 - N / 64)
        h = (rng.random() * 2 - 1) * np.pi / 2
        phase += h * np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * sig * sig))
    return np.asfortranarray(amp * np.exp(1j * phase))


def image_model_numpy(O: np.ndarray, P: np.ndarray, led: np.ndarray, pratio: int) -> np.ndarray:
    """|IDFT_n(ifftshift(crop(DFT_N O) . P))| / (n^2 r^2) for every LED (float64),
    the amplitude the solve compares with Y (cuFPM_v3/cuda/kernels.cuh:19-65,223)."""
    Nr, Nc = O.shape
    nr, nc = P.shape
    S = np.fft.fft2(O)
    out = np.empty((nr, nc, led.shape[1]))
    ii = np.arange(nr)[:, None]
    jj = np.arange(nc)[None, :]
    for l in range(led.shape[1]):
        kl, kt = int(led[0, l]), int(led[1, l])
        rows = (ii + kt - 1 + Nr // 2) % Nr
        cols = (jj + kl - 1 + Nc // 2) % Nc
        sub = S[rows, cols] * P
        z = np.fft.ifft2(np.fft.ifftshift(sub)) * (nr * nc)
        out[:, :, l] = np.abs(z) / (nr * nc * pratio * pratio)
    return out


def fortran_empty(shape, dtype, device):
    """torch tensor laid out column-major for MATLAB shape `shape`: returned as
    the row-major buffer of reversed shape (element (i, j[, l]) of the MATLAB
    array is t[l, j, i])."""
    import torch

    return torch.empty(tuple(reversed(shape)), dtype=dtype, device=device)


def make_device(n: int, pratio: int, seed: int = 20261203, device: str = "cuda", led_grid: int = LED_GRID,
                defocus_rad: float = 0.5, keep_object: bool = False, pupil_coeffs=None):
    """The same generator on the GPU for large workloads (torch tensors in
    column-major layout, transforms and the image model through libfpm_amd).
    Returns a dict of device tensors + host LED/mask arrays.  keep_object
    adds the true object in the solve's scale (O_true_dev = kappa O, where
    kappa maps the object to the prepared amplitudes sqrt((I - min) /
    (max - min))) and its phase field (phase_true_dev).  pupil_coeffs (5
    Zernike coefficients, rad) replaces the defocus-only true pupil; the
    dict's "P_true_dev" holds the true pupil in the device layout."""
    import torch
    import torch.nn.functional as F

    from . import _check, lib

    N = n * pratio
    dev = torch.device(device)
    stream = torch.cuda.current_stream(dev).cuda_stream
    led, _, nai = led_offsets(n, pratio, wavelength_um=WAVELENGTH_UM, na=NA, pixel_um=PIXEL_UM,
                              led_pitch_um=LED_PITCH_UM, led_height_um=LED_HEIGHT_UM, led_grid=led_grid)
    L = led.shape[1]
    A, P_true = pupil(n, defocus_rad)
    if pupil_coeffs is not None:
        P_true = zernike_pupil(n, pupil_coeffs, A)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    # amplitude: band-limited noise (the filter is even, so the layout
    # transpose of the torch view does not matter)
    noise = torch.rand((N, N), generator=gen, device=dev, dtype=torch.float32).to(torch.complex64)
    spec = torch.empty_like(noise)
    _check(lib().fpm_fft2(noise.data_ptr(), spec.data_ptr(), N, N, 1, -1, 1, stream))
    f = torch.fft.fftfreq(N, d=PIXEL_UM / pratio, device=dev, dtype=torch.float64) * 2 * np.pi
    cut = (NA + float(nai.max())) * 2 * np.pi / WAVELENGTH_UM
    mask = ((f[:, None] ** 2 + f[None, :] ** 2) <= cut * cut).to(torch.float32)
    spec *= mask
    del mask
    _check(lib().fpm_fft2(spec.data_ptr(), noise.data_ptr(), N, N, 1, 1, 1, stream))
    amp = noise.real.clone()
    del noise, spec
    amp = 0.3 + 0.7 * (amp - amp.min()) / (amp.max() - amp.min())
    # phase: 20 Gaussians (separable outer products)
    phase = torch.zeros((N, N), device=dev, dtype=torch.float32)
    g = np.random.Generator(np.random.Philox(seed))
    ax = torch.arange(N, device=dev, dtype=torch.float32)
    for _ in range(20):
        cy, cx = g.random(2) * N
        sig = N / 64 + g.random() * (N / 16 - N / 64)
        h = (g.random() * 2 - 1) * np.pi / 2
        gy = torch.exp(-((ax - float(cy)) ** 2) / (2 * sig * sig))
        gx = torch.exp(-((ax - float(cx)) ** 2) / (2 * sig * sig))
        phase.add_(h * torch.outer(gx, gy))
    O_true = torch.polar(amp, phase)
    phase_keep = phase if keep_object else None
    del amp, phase
    Pt = torch.from_numpy(np.ascontiguousarray(P_true.T)).to(dev)
    ledd = torch.from_numpy(np.ascontiguousarray(led.T)).to(dev)
    I = fortran_empty((n, n, L), torch.float32, dev)
    _check(lib().fpm_forward_intensity(O_true.data_ptr(), N, N, Pt.data_ptr(), n, n, L, ledd.data_ptr(),
                                       float(pratio), 1, I.data_ptr(), 1, stream))
    if not keep_object:
        del O_true
    scale = PEAK_COUNTS / float(I[0].max())
    I *= scale
    I = torch.poisson(I, generator=gen)
    I += OFFSET_COUNTS
    counts = I.round_().clamp_(0, 65535).to(torch.int32)
    del I
    mn, mx = int(counts.min()), int(counts.max())
    centre = torch.sqrt((counts[0].to(torch.float32) - mn) / float(mx - mn))
    raw = counts.to(torch.uint16)
    del counts
    O0 = F.interpolate(centre[None, None], scale_factor=pratio, mode="bilinear", align_corners=False)[0, 0]
    O0 = O0.to(torch.complex64).contiguous()
    P0 = torch.from_numpy(np.ascontiguousarray(A.T.astype(np.complex64))).to(dev)
    Ad = torch.from_numpy(np.ascontiguousarray(A.T)).to(dev)
    out = {"n": n, "N": N, "L": L, "pratio": pratio, "led": led, "led_dev": ledd, "A": A, "A_dev": Ad,
           "raw_dev": raw, "O0_dev": O0, "P0_dev": P0, "P_true_dev": Pt, "na_illum": nai, "raw_min": mn,
           "raw_max": mx}
    if keep_object:
        kappa = float(np.sqrt(scale / float(mx - mn)))
        out["O_true_dev"] = O_true * kappa
        out["phase_true_dev"] = phase_keep
        out["kappa"] = kappa
    torch.cuda.synchronize(dev)
    return out


def make_numpy(n: int, pratio: int, seed: int = 20261203, led_grid: int = LED_GRID, defocus_rad: float = 0.5,
               noise: bool = True) -> Dataset:
    """Whole dataset on the host (small and medium sizes)."""
    N = n * pratio
    led, _, nai = led_offsets(n, pratio, wavelength_um=WAVELENGTH_UM, na=NA, pixel_um=PIXEL_UM,
                              led_pitch_um=LED_PITCH_UM, led_height_um=LED_HEIGHT_UM, led_grid=led_grid)
    A, P_true = pupil(n, defocus_rad)
    O = object_numpy(N, pratio, float(nai.max()), seed)
    amp = image_model_numpy(O, P_true, led, pratio)
    I = amp * amp
    I = I * (PEAK_COUNTS / I[:, :, 0].max())
    rng = np.random.Generator(np.random.Philox(seed + 1))
    if noise:
        I = rng.poisson(I).astype(np.float64)
    raw = np.clip(np.rint(I + OFFSET_COUNTS), 0, 65535).astype(np.uint16)
    raw = np.asfortranarray(raw)
    Y = raw.astype(np.float32) - np.float32(raw.min())
    Y = np.sqrt(Y / np.float32(Y.max())).astype(np.float32)
    O0 = bilinear_upsample(Y[:, :, 0].astype(np.float64), pratio).astype(np.complex64)
    return Dataset(n=n, pratio=pratio, led=led, A=A, P_true=P_true, O_true=O.astype(np.complex64), raw=raw,
                   amp=np.asfortranarray(Y), O0=np.asfortranarray(O0), P0=np.asfortranarray(A.astype(np.complex64)),
                   na_illum=nai)
