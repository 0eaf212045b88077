"""Pins of the CPU oracle (no GPU): each check is chosen so that a plausible
mistake (dropped term, wrong sign/index, conj, scale) fails it."""
import numpy as np
import pytest

import oracle
from fpm_amd import synth
from helpers import margin_problem, rel_l2

RNG = np.random.default_rng(1234)


def _cplx(shape, rng=RNG):
    return np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


@pytest.mark.parametrize("rows,cols,batch", [(16, 16, 1), (64, 32, 3), (32, 128, 2), (8, 512, 1),
                                             (192, 192, 2), (48, 96, 1), (12, 24, 2), (768, 36, 1)])
def test_oracle_fft_matches_numpy(rows, cols, batch):
    x = _cplx((rows, cols, batch))
    ref = np.fft.fft2(x, axes=(0, 1))
    assert np.abs(oracle.fft2(x) - ref).max() < 1e-10 * np.abs(ref).max()
    inv = np.fft.ifft2(x, axes=(0, 1)) * rows * cols
    assert np.abs(oracle.fft2(x, inverse=True) - inv).max() < 1e-10 * np.abs(inv).max()
    assert rel_l2(oracle.fft2(x, prec="f32"), ref) < 1e-6


@pytest.mark.parametrize("n", [16, 12, 18])
def test_oracle_fft_brute_force_dft(n):
    """Powers of two and the mixed-radix lengths 2^a 3^b (radices 4, 2, 3)."""
    x = _cplx((n, n))
    k = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(k, k) / n)
    ref = W @ x @ W.T  # DFT along both axes
    assert np.abs(oracle.fft2(x) - ref).max() < 1e-11


def test_fft_golden_fixture():
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_vectors.npz"))
    assert np.abs(oracle.fft2(d["x2d"]) - d["X2d"]).max() < 1e-9 * np.abs(d["X2d"]).max()


def test_fft_properties():
    n = 32
    x = _cplx((n, n))
    X = oracle.fft2(x)
    # Parseval
    assert abs(np.vdot(X, X).real - n * n * np.vdot(x, x).real) < 1e-9 * n * n * np.vdot(x, x).real
    # delta <-> constant
    d = np.zeros((n, n), complex, order="F")
    d[0, 0] = 1
    assert np.allclose(oracle.fft2(d), 1)
    # shift theorem: circular shift by (a, b) multiplies by a phase ramp
    a, b = 3, 5
    xs = np.asfortranarray(np.roll(x, (a, b), axis=(0, 1)))
    k = np.arange(n)
    ramp = np.exp(-2j * np.pi * (np.outer(k, np.ones(n)) * a + np.outer(np.ones(n), k) * b) / n)
    assert np.abs(oracle.fft2(xs) - X * ramp).max() < 1e-9 * np.abs(X).max()


@pytest.mark.parametrize("M", [4, 6, 8, 16, 32])
def test_fftshift_index_closed_form(M):
    # the folded shift (k + M/2) mod M of kernels.cuh:45-47,62-64 is numpy's
    # fftshift and, for even M, its own inverse
    v = np.arange(M)
    idx = (v + M // 2) % M
    assert (np.fft.fftshift(v) == v[idx]).all()
    assert (np.fft.ifftshift(v) == v[idx]).all()


def test_gradient_matches_finite_differences():
    n, r = 16, 2
    O, P, Y, led = margin_problem(n, r)
    GO, GP, loss, margin = oracle.gradient(O, P, Y, led, r)
    assert margin > 5e-3
    rng = np.random.default_rng(3)
    h = 1e-4
    Od = O.astype(np.complex128)
    Pd = P.astype(np.complex128)
    for _ in range(2):
        d = _cplx(O.shape, rng)
        fd = (oracle.loss(Od + h * d, Pd, Y, led, r) - oracle.loss(Od - h * d, Pd, Y, led, r)) / (2 * h)
        # G_O = r^2 (d/dRe + i d/dIm) L
        assert abs(fd - np.real(np.vdot(GO, d)) / r ** 2) < 1e-6 * max(1.0, abs(fd))
        dp = _cplx(P.shape, rng)
        fd = (oracle.loss(Od, Pd + h * dp, Y, led, r) - oracle.loss(Od, Pd - h * dp, Y, led, r)) / (2 * h)
        # G_P = n^2 r^4 (d/dRe + i d/dIm) L
        assert abs(fd - np.real(np.vdot(GP, dp)) / (n * n * r ** 4)) < 1e-6 * max(1.0, abs(fd))


def test_loss_matches_numpy_restatement():
    n, r = 16, 2
    O, P, Y, led = margin_problem(n, r)
    model = synth.image_model_numpy(O.astype(np.complex128), P.astype(np.complex128), led, r)
    e = model - Y
    ref = (np.abs(np.roll(e, -1, axis=1) - e) + np.abs(np.roll(e, -1, axis=0) - e)).sum()
    assert abs(oracle.loss(O, P, Y, led, r) - ref) < 1e-9 * ref


def test_gradient_f32_close_to_f64():
    O, P, Y, led = margin_problem(16, 2)
    g64 = oracle.gradient(O, P, Y, led, 2, prec="f64")
    g32 = oracle.gradient(O, P, Y, led, 2, prec="f32")
    assert rel_l2(g32[0], g64[0]) < 1e-5
    assert rel_l2(g32[1], g64[1]) < 1e-5


def test_solve_update_rule_scalar_restatement():
    """One step of the oracle solve equals an elementwise restatement of
    kernels.cuh:364-410 applied to the oracle gradient."""
    n, r = 16, 2
    O, P, Y, led = margin_problem(n, r)
    A, _ = synth.pupil(n)
    lr, beta, eps = 5e-3, 0.999, 1e-4
    L = led.shape[1]
    O1, P1, _ = oracle.solve(O, P, Y, led, A, r, 1, L, (lr, beta, eps), prec="f64")
    GO, GP, _, _ = oracle.gradient(O, P, Y, led, r)
    # the mex casts lr, beta, eps to float (cuFPM_pure_v3.cu:63-65)
    lr, beta, eps = (float(np.float32(v)) for v in (lr, beta, eps))
    m = (1 - beta) * np.abs(GO) ** 2
    Oref = O - lr * GO / (np.sqrt(m) + eps)
    assert rel_l2(O1, Oref) < 1e-12
    mp = (1 - beta) * np.abs(GP) ** 2
    w = P - lr * GP / (np.sqrt(mp) + eps)
    a = A.astype(np.float64)
    ab = a * np.maximum(np.minimum(np.abs(w), a + 0.2), a - 0.2)
    Pref = ab * np.exp(1j * np.angle(w))
    assert rel_l2(P1, Pref) < 1e-12
    # invariant: |P| in A [A - 0.2, A + 0.2]
    mag = np.abs(P1)
    assert (mag <= a * (a + 0.2) + 1e-12).all() and (mag >= a * (a - 0.2) - 1e-12).all()


def test_lr_schedule_halves_after_epoch_12():
    """With zero residual curvature the per-step displacement is lr-scaled:
    run 16 epochs and check the loss trajectory is finite and the lr halving
    is visible as a shrinking step (epochs 14, 15 move less than 13)."""
    n, r = 16, 2
    O, P, Y, led = margin_problem(n, r)
    A, _ = synth.pupil(n)
    L = led.shape[1]
    prev = O.astype(np.complex128)
    steps = []
    for ep in range(1, 17):
        Oe, _, _ = oracle.solve(O, P, Y, led, A, r, ep, L, (1e-3, 0.999, 1e-4), prec="f64")
        steps.append(np.linalg.norm(Oe - prev))
        prev = Oe
    assert steps[14] < 0.75 * steps[12] and steps[15] < 0.75 * steps[13]


def test_prepare_amplitude_oracle():
    raw = RNG.integers(100, 40000, size=(16, 16, 4)).astype(np.uint16)
    got = oracle.prepare_amplitude(raw)
    f = raw.astype(np.float32) - np.float32(raw.min())
    ref = np.sqrt(f / np.float32(f.max()))
    assert np.array_equal(got, ref)


def test_led_offsets_closed_forms():
    n, r = 128, 4
    led, k, margin = oracle.led_offsets_matlab(n, r)
    N = n * r
    # centre LED (first in spiral order): window centred on DC
    assert led[0, 0] == N // 2 - n // 2 + 1 and led[1, 0] == N // 2 - n // 2 + 1
    assert k[0] == (19 * 19 + 1) // 2
    assert margin > 1e-3
    # mirror symmetry of the grid: LED (i, j) and (20-i, 20-j) are mirrored
    # around the centre window start
    c = N // 2 - n // 2 + 1
    pos = {int(kk): (int(led[0, q]), int(led[1, q])) for q, kk in enumerate(k)}
    for kk, (kl, kt) in pos.items():
        i, j = (kk - 1) // 19 + 1, (kk - 1) % 19 + 1
        mk = (20 - j) + (20 - i - 1) * 19
        mkl, mkt = pos[mk]
        assert kl - c == -(mkl - c) and kt - c == -(mkt - c)
    assert sorted(k.tolist()) == list(range(1, 362))
