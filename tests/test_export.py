"""Result export (images + per-epoch log), host only."""
import json

import numpy as np
from PIL import Image

from fpm_amd import export


def test_mat2gray_and_conversions():
    x = np.array([[1.0, 3.0], [2.0, 5.0]])
    g = export.mat2gray(x)
    assert g.min() == 0 and g.max() == 1 and abs(g[0, 1] - 0.5) < 1e-15
    assert export.to_uint8(np.array([0.0, 0.5, 1.0])).tolist() == [0, 128, 255]
    assert export.to_uint16(np.array([0.0, 0.5, 1.0])).tolist() == [0, 32768, 65535]
    assert (export.mat2gray(np.ones((3, 3))) == 0).all()


def test_export_result_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    O = (rng.random((32, 32)) + 0.1) * np.exp(1j * rng.random((32, 32)))
    P = np.exp(1j * rng.random((16, 16)))
    paths = export.export_result(O, P, str(tmp_path / "green"), raw_centre=rng.random((16, 16)))
    assert len(paths) == 5
    amp = np.array(Image.open(tmp_path / "green_amp.tif"))
    assert amp.dtype == np.uint16 and amp.shape == (32, 32)
    assert np.array_equal(amp, export.to_uint16(export.mat2gray(np.abs(O))))
    out = np.array(Image.open(tmp_path / "green_out.png"))
    assert np.array_equal(out, export.to_uint8(export.mat2gray(np.abs(O) ** 2)))


def test_psnr_ssim_basic():
    rng = np.random.default_rng(1)
    ref = rng.random((64, 64))
    assert export.psnr(ref, ref) == float("inf")
    assert abs(export.ssim(ref, ref) - 1.0) < 1e-12
    noisy = ref + 0.01 * rng.standard_normal(ref.shape)
    p = export.psnr(noisy, ref)
    assert 35 < p < 45  # 10 log10(1 / 1e-4) ~ 40 dB
    assert export.ssim(noisy, ref) < 1.0


def test_psnr_ssim_use_matlabs_range_for_floating_images():
    """MATLAB's psnr(A, ref) and ssim(A, ref) on single/double images take
    the class range 1 as peak / dynamic range, whatever the images' own
    maximum: an offset of 0.1 on an image in [0, 0.5] is 20 dB exactly
    (10 log10(1 / 0.01)), not 10 log10(0.25 / 0.01)."""
    rng = np.random.default_rng(3)
    ref = 0.5 * rng.random((64, 64))
    assert abs(export.psnr(ref + 0.1, ref) - 20.0) < 1e-9
    assert abs(export.psnr(ref + 0.1, ref, peak=0.5) - 10 * np.log10(0.25 / 0.01)) < 1e-9
    x = ref + 0.05 * rng.standard_normal(ref.shape)
    assert export.ssim(x, ref) == export.ssim(x, ref, data_range=1.0)
    assert export.ssim(x, ref) != export.ssim(x, ref, data_range=float(ref.max() - ref.min()))


def test_epoch_log_save(tmp_path):
    log = export.EpochLog()
    rng = np.random.default_rng(2)
    t = rng.random((16, 16))
    for e in range(3):
        log.add(0.1 * (e + 1), loss=10.0 - e, amp=t + 0.01 * e, truth=t)
    mat, js = log.save(str(tmp_path / "log"))
    from scipy.io import loadmat

    d = loadmat(mat)
    assert np.allclose(d["time_data"].ravel(), [0.1, 0.2, 0.3])
    assert json.load(open(js))["loss_data"] == [10.0, 9.0, 8.0]
