"""Split-precision FFT engine (opts.fft_engine = FPM_FFT_SPLIT_F16): the
radix-16 DFT stages as FP16 hi+lo MFMA products with FP32 accumulation.
Every result is compared with FP64 next to the FP32 engine's error on the
same input; the bounds are stated per test."""
import numpy as np
import pytest

import fpm_amd
import oracle
from fpm_amd import synth
from helpers import margin_problem, rel_l2, small_leds

pytestmark = pytest.mark.gpu
SPLIT = fpm_amd.FFT_SPLIT_F16


def _cplx(shape, seed):
    rng = np.random.default_rng(seed)
    return np.asfortranarray((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64))


@pytest.mark.parametrize("rows,cols,batch", [
    (16, 16, 1), (32, 64, 3), (512, 512, 2),      # single strided pass
    (64, 1024, 1), (16, 2048, 2), (32, 4096, 1),  # four-step columns
    (1024, 16, 2), (16384, 16, 1), (16, 16384, 1), (2048, 2048, 1),
])
def test_split_fft2_matches_fp64(rows, cols, batch):
    """Forward and inverse transforms vs numpy in FP64: the split engine
    within 4e-6 rel-L2 (hi+lo keeps 22 significant bits per operand; the
    FP32 engine's own error is printed beside it)."""
    x = _cplx((rows, cols, batch), rows * 7 + cols)
    ref = np.fft.fft2(x.astype(np.complex128), axes=(0, 1))
    e32 = rel_l2(fpm_amd.fft2(x), ref)
    got = fpm_amd.fft2(x, engine=SPLIT)
    e16 = rel_l2(got, ref)
    inv = fpm_amd.fft2(got, inverse=True, engine=SPLIT) / (rows * cols)
    ei = rel_l2(inv, x)
    print(f"[split-fft] {rows}x{cols}x{batch}: rel-L2 vs FP64 split {e16:.2e} (FP32 engine {e32:.2e}), "
          f"round trip {ei:.2e}", flush=True)
    assert e16 < 4e-6
    assert ei < 6e-6


def test_split_fft2_wide_dynamic_range():
    """A spectrum-like input (one huge DC term, 1e-6 relative detail): the
    per-butterfly power-of-two scaling keeps the error relative to each
    butterfly's own magnitude, so rel-L2 stays at the random-input level."""
    x = _cplx((256, 256, 1), 3) * np.float32(1e-3)
    x[0, 0, 0] = 1e5
    x[5, 7, 0] = -3e4
    ref = np.fft.fft2(x.astype(np.complex128), axes=(0, 1))
    got = fpm_amd.fft2(x, engine=SPLIT)
    assert rel_l2(got, ref) < 4e-6
    # the small detail survives: error of the transform of the detail alone
    d = x.copy()
    d[0, 0, 0] = 0
    d[5, 7, 0] = 0
    ref_d = np.fft.fft2(d.astype(np.complex128), axes=(0, 1))
    assert rel_l2(fpm_amd.fft2(d, engine=SPLIT), ref_d) < 4e-6


def test_split_fft2_zero_and_delta():
    """Zero input stays exactly zero (the scale of an all-zero butterfly is
    finite); a delta transforms to a constant to FP32 precision."""
    z = np.zeros((64, 64), np.complex64, order="F")
    assert not np.any(fpm_amd.fft2(z, engine=SPLIT))
    d = z.copy()
    d[0, 0] = 1.0
    got = fpm_amd.fft2(d, engine=SPLIT)
    assert np.max(np.abs(got - 1.0)) < 1e-6


@pytest.mark.parametrize("n,r,grid", [(32, 4, 5), (64, 16, 3), (128, 2, 3), (1024, 2, 3)])
def test_split_gradient_vs_fp64_oracle(n, r, grid):
    """One-minibatch gradient with every transform on the split engine vs
    the FP64 oracle: rel-L2 <= 1e-5, the bar of the FP32 form."""
    O, P, Y, led = margin_problem(n, r, grid)
    GO, GP, loss, _ = oracle.gradient(O, P, Y, led, r)
    gO32, gP32, _ = fpm_amd.fused_gradient(O, P, Y, led, r)
    gO, gP, gl = fpm_amd.fused_gradient(O, P, Y, led, r, engine=SPLIT)
    print(f"[split-grad] n={n} r={r}: G_O {rel_l2(gO, GO):.2e} (FP32 {rel_l2(gO32, GO):.2e}), "
          f"G_P {rel_l2(gP, GP):.2e} (FP32 {rel_l2(gP32, GP):.2e})", flush=True)
    assert rel_l2(gO, GO) <= 1e-5
    assert rel_l2(gP, GP) <= 1e-5
    assert abs(gl - loss) <= 1e-5 * loss


def test_split_solve_20_epochs_matches_fp32_oracle():
    """20-epoch solve on the split engine vs the FP32 oracle (same init and
    batch partition): rel-L2 <= 1e-3 on O and P, like the FP32 form."""
    n, r, grid = 32, 4, 7
    led = small_leds(n, r, grid)
    A, Ptrue = synth.pupil(n)
    O = synth.object_numpy(n * r, r, 0.3, 11)
    amp = synth.image_model_numpy(O, Ptrue.astype(np.complex128), led, r).astype(np.float32)
    O0 = np.asfortranarray(synth.bilinear_upsample(amp[:, :, 0].astype(np.float64), r).astype(np.complex64))
    P0 = np.asfortranarray(A.astype(np.complex64))
    amp = np.asfortranarray(amp)
    opt = (5e-4, 0.999, 1e-4)
    Os, Ps, ls = fpm_amd.cuFPM_pure_v3(O0, P0, amp, led, A, r, 20, 10, opt, loss=True, engine=SPLIT)
    Oo, Po, lo = oracle.solve(O0, P0, amp, led, A, r, 20, 10, opt, prec="f32")
    print(f"[split-solve] rel-L2 vs FP32 oracle O {rel_l2(Os, Oo):.2e} P {rel_l2(Ps, Po):.2e}", flush=True)
    assert rel_l2(Os, Oo) <= 1e-3
    assert rel_l2(Ps, Po) <= 1e-3
    assert np.allclose(ls, lo, rtol=1e-3)


def test_split_engine_with_graph_replay_and_atomic_stitch():
    """The split engine composes with the other options: graph replay gives
    the same bits as eager launches, and the atomic stitch stays within the
    20-epoch bar of the FP32 oracle."""
    n, r, grid = 32, 4, 7
    led = small_leds(n, r, grid)
    A, Ptrue = synth.pupil(n)
    O = synth.object_numpy(n * r, r, 0.3, 12)
    amp = np.asfortranarray(synth.image_model_numpy(O, Ptrue.astype(np.complex128), led, r).astype(np.float32))
    O0 = np.asfortranarray(synth.bilinear_upsample(amp[:, :, 0].astype(np.float64), r).astype(np.complex64))
    P0 = np.asfortranarray(A.astype(np.complex64))
    opt = (5e-4, 0.999, 1e-4)
    a = fpm_amd.cuFPM_pure_v3(O0, P0, amp, led, A, r, 15, 10, opt, loss=True, engine=SPLIT)
    b = fpm_amd.cuFPM_pure_v3(O0, P0, amp, led, A, r, 15, 10, opt, loss=True, engine=SPLIT, graph=True)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    Oa, Pa = fpm_amd.cuFPM_pure_v3(O0, P0, amp, led, A, r, 20, 10, opt, engine=SPLIT, stitch=fpm_amd.STITCH_ATOMIC)
    Oo, Po, _ = oracle.solve(O0, P0, amp, led, A, r, 20, 10, opt, prec="f32")
    assert rel_l2(Oa, Oo) <= 1e-3 and rel_l2(Pa, Po) <= 1e-3
