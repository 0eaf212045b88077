"""Result export: images of the reconstruction and a per-epoch quality/time log.

Follows what the reference's drivers write after a solve:
* ``imwrite(mat2gray(abs(w1).^2), '<colour>_out.png')`` and the raw centre image
  (codes/main_run_FPM.m:174-175): 8-bit PNG of min-max scaled intensity;
* ``imwrite(uint16((2^16-1) * mat2gray(abs(w1))), ...tif)`` (benchmark/cuFPM_recon_2025_circular.m:222);
* ``save(..., 'ssim_data', 'psnr_data', 'time_data')`` per epoch (benchmark/cuFPM_recon_2025_circular.m:189-228).
"""
from __future__ import annotations

import json
import os
from typing import List, Optional

import numpy as np


def mat2gray(x: np.ndarray) -> np.ndarray:
    """Min-max scale to [0, 1] (MATLAB mat2gray with default limits)."""
    x = np.asarray(x, dtype=np.float64)
    lo, hi = float(x.min()), float(x.max())
    if hi == lo:
        return np.zeros_like(x)
    return (x - lo) / (hi - lo)


def to_uint8(g: np.ndarray) -> np.ndarray:
    """im2uint8 of a [0, 1] double image: round(255 x)."""
    return np.clip(np.rint(g * 255.0), 0, 255).astype(np.uint8)


def to_uint16(g: np.ndarray) -> np.ndarray:
    """uint16((2^16 - 1) x) with MATLAB's round-half-away conversion."""
    v = g * 65535.0
    return np.clip(np.floor(v + 0.5), 0, 65535).astype(np.uint16)


def write_png8(path: str, img: np.ndarray) -> None:
    from PIL import Image

    Image.fromarray(to_uint8(mat2gray(img))).save(path)


def write_tiff16(path: str, img: np.ndarray) -> None:
    from PIL import Image

    Image.fromarray(to_uint16(mat2gray(img))).save(path)


def export_result(O: np.ndarray, P: np.ndarray, prefix: str, raw_centre: Optional[np.ndarray] = None) -> List[str]:
    """Writes <prefix>_out.png (|O|^2), <prefix>_amp.tif (|O|, 16 bit),
    <prefix>_phase.png (angle O), <prefix>_pupil_phase.png and, if given,
    <prefix>_raw.png (centre image^2).  Returns the paths."""
    paths = []
    O = np.asarray(O)
    P = np.asarray(P)
    items = [("_out.png", np.abs(O) ** 2, write_png8), ("_amp.tif", np.abs(O), write_tiff16),
             ("_phase.png", np.angle(O), write_png8), ("_pupil_phase.png", np.angle(P), write_png8)]
    if raw_centre is not None:
        items.append(("_raw.png", np.asarray(raw_centre, dtype=np.float64) ** 2, write_png8))
    for suffix, img, fn in items:
        p = prefix + suffix
        fn(p, img)
        paths.append(p)
    return paths


def psnr(x: np.ndarray, ref: np.ndarray, peak: Optional[float] = None) -> float:
    """Peak signal-to-noise ratio in dB.  peak defaults to 1, the range MATLAB's
    psnr(A, ref) takes for single or double images, as the reference's log
    calls it (benchmark/cuFPM_recon_2025_circular.m:193: psnr(gt,
    single(abs(w2))))."""
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    mse = float(np.mean((x - ref) ** 2))
    pk = 1.0 if peak is None else peak
    return float("inf") if mse == 0 else 10.0 * np.log10(pk * pk / mse)


def ssim(x: np.ndarray, ref: np.ndarray, data_range: Optional[float] = None) -> float:
    """Mean SSIM with an 11x11 Gaussian window (sigma 1.5, replicated
    borders), K1 0.01, K2 0.03.  data_range defaults to 1, the range MATLAB's
    ssim takes for single or double images (benchmark/cuFPM_recon_2025_
    circular.m:194)."""
    from scipy.ndimage import gaussian_filter

    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(ref, dtype=np.float64)
    L = 1.0 if data_range is None else data_range
    c1, c2 = (0.01 * L) ** 2, (0.03 * L) ** 2
    f = lambda a: gaussian_filter(a, 1.5, truncate=3.5, mode="nearest")  # 11-tap window
    mx, my = f(x), f(y)
    sxx, syy, sxy = f(x * x) - mx * mx, f(y * y) - my * my, f(x * y) - mx * my
    s = ((2 * mx * my + c1) * (2 * sxy + c2)) / ((mx * mx + my * my + c1) * (sxx + syy + c2))
    return float(s.mean())


class EpochLog:
    """Per-epoch time / loss / PSNR / SSIM, saved as .mat (scipy) and JSON."""

    def __init__(self):
        self.time_data: List[float] = []
        self.loss_data: List[float] = []
        self.psnr_data: List[float] = []
        self.ssim_data: List[float] = []

    def add(self, seconds: float, loss: float = float("nan"), amp: Optional[np.ndarray] = None,
            truth: Optional[np.ndarray] = None) -> None:
        self.time_data.append(float(seconds))
        self.loss_data.append(float(loss))
        if amp is not None and truth is not None:
            self.psnr_data.append(psnr(amp, truth))
            self.ssim_data.append(ssim(amp, truth))
        else:
            self.psnr_data.append(0.0)
            self.ssim_data.append(0.0)

    def save(self, path_noext: str) -> List[str]:
        from scipy.io import savemat

        d = {"time_data": np.array(self.time_data), "loss_data": np.array(self.loss_data),
             "psnr_data": np.array(self.psnr_data), "ssim_data": np.array(self.ssim_data)}
        savemat(path_noext + ".mat", d)
        with open(path_noext + ".json", "w") as fh:
            json.dump({k: v.tolist() for k, v in d.items()}, fh)
        return [path_noext + ".mat", path_noext + ".json"]
