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
                                             (192, 192, 2), (48, 96, 1), (12, 24, 2), (768, 36, 1),
                                             (200, 40, 1), (14, 98, 2), (134, 22, 1), (254, 12, 1), (12288, 4, 1)])
def test_oracle_fft_matches_numpy(rows, cols, batch):
    x = _cplx((rows, cols, batch))
    ref = np.fft.fft2(x, axes=(0, 1))
    assert np.abs(oracle.fft2(x) - ref).max() < 1e-10 * np.abs(ref).max()
    inv = np.fft.ifft2(x, axes=(0, 1)) * rows * cols
    assert np.abs(oracle.fft2(x, inverse=True) - inv).max() < 1e-10 * np.abs(inv).max()
    assert rel_l2(oracle.fft2(x, prec="f32"), ref) < 1e-6


@pytest.mark.parametrize("n", [16, 12, 18, 10, 14, 22, 26, 34])
def test_oracle_fft_brute_force_dft(n):
    """Powers of two and mixed-radix lengths (radices 4, 2, 3, 5, 7 and the
    primes 11, 13, 17 as direct sums)."""
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


def _state_problem():
    """Margin inputs, one minibatch per epoch (batch = L): every sign of the
    loss stencil is decided by >= 0.01, so with a tiny learning rate the
    gradient stays constant to ~1e-5 relative over 16 steps."""
    n, r = 16, 2
    O, P, Y, led = margin_problem(n, r)
    A, _ = synth.pupil(n)
    return n, r, O, P, Y, led, A


def _closed_form_step(g, lr_e, beta, k, eps):
    """RMSprop without bias correction from m_0 = 0 under a constant
    gradient g: m_k = (1 - beta^k) |g|^2 exactly, so step k moves the
    variable by lr_e g / (sqrt(1 - beta^k) |g| + eps).  This is the
    recurrence of kernels.cuh:368-377 summed in closed form; it fails if
    the m store is dropped (m_k = (1 - beta)|g|^2: steps stay equal) or if
    the carried term is mis-scaled."""
    return lr_e * g / (np.sqrt(1.0 - beta ** k) * np.abs(g) + eps)


@pytest.mark.parametrize("prec,lr,tol", [("f64", 1e-7, 2e-4), ("f32", 3e-6, 5e-3)])
def test_solve_carried_moments_closed_form(prec, lr, tol):
    """Steps 1..4 of the oracle solve against the closed form above, for O
    and for P (inside the pupil support |P| stays in A[A-0.2, A+0.2], so the
    clamp is the identity and the same closed form holds).  Step k+1 / step
    k is sqrt((1 - beta^k) / (1 - beta^(k+1))), 0.707 for k = 1; a solve
    that forgot m between steps gives 1."""
    n, r, O, P, Y, led, A = _state_problem()
    L = led.shape[1]
    beta = float(np.float32(0.999))
    eps = float(np.float32(1e-4))
    lr32 = float(np.float32(lr))
    GO, GP, _, _ = oracle.gradient(O, P, Y, led, r)          # pinned by finite differences
    prevO, prevP = O.astype(np.complex128), P.astype(np.complex128)
    inside = A > 0
    for k in range(1, 5):
        Ok, Pk, _ = oracle.solve(O, P, Y, led, A, r, k, L, (lr, 0.999, 1e-4), prec=prec)
        dO, dP = prevO - Ok, prevP - Pk
        assert rel_l2(dO, _closed_form_step(GO, lr32, beta, k, eps)) < tol, k
        assert rel_l2(dP[inside], _closed_form_step(GP[inside], lr32, beta, k, eps)) < tol, k
        assert np.abs(Pk[~inside]).max() == 0.0
        prevO, prevP = Ok.astype(np.complex128), Pk.astype(np.complex128)
        if k == 2:
            ratio = np.linalg.norm(dO) / step1
            assert abs(ratio - np.sqrt((1 - beta) / (1 - beta ** 2))) < 5 * tol
        step1 = np.linalg.norm(dO)


def test_lr_halving_epoch_exact():
    """cuFPM_pure_v3.cu:204-206 halves lr at the end of every epoch with
    index > 12: epochs 0..13 step with lr, epoch 14 with lr/2, epoch 15 with
    lr/4.  With one minibatch per epoch, each epoch's displacement is
    checked against the closed form with that epoch's lr, so halving one
    epoch early (after epoch 12) or late (after 14) fails."""
    n, r, O, P, Y, led, A = _state_problem()
    L = led.shape[1]
    lr = 1e-7
    beta = float(np.float32(0.999))
    eps = float(np.float32(1e-4))
    lr32 = float(np.float32(lr))
    GO, _, _, _ = oracle.gradient(O, P, Y, led, r)
    prev = None
    disp = {}
    for ep in range(11, 17):  # solve of `ep` epochs ends after epoch index ep-1
        Oe, _, _ = oracle.solve(O, P, Y, led, A, r, ep, L, (lr, 0.999, 1e-4), prec="f64")
        if prev is not None:
            e = ep - 1  # displacement of epoch index e
            lr_e = lr32 * (0.5 ** max(0, e - 13))
            d = prev - Oe
            assert rel_l2(d, _closed_form_step(GO, lr_e, beta, e + 1, eps)) < 2e-4, e
            disp[e] = np.linalg.norm(d)
        prev = Oe.astype(np.complex128)
    f = lambda e: np.sqrt(1 - beta ** (e + 1))  # noqa: E731
    assert abs(disp[12] / disp[11] - f(11) / f(12)) < 1e-3   # no halving before epoch 13
    assert abs(disp[13] / disp[12] - f(12) / f(13)) < 1e-3   # epoch 13 still at full lr
    assert abs(disp[14] / disp[13] - 0.5 * f(13) / f(14)) < 1e-3
    assert abs(disp[15] / disp[14] - 0.5 * f(14) / f(15)) < 1e-3


def test_lr_halve_from_option():
    """The schedule option moves the first halved epoch: halve_from = h gives
    lr/2 from epoch h+1 on (13 is the reference)."""
    n, r, O, P, Y, led, A = _state_problem()
    L = led.shape[1]
    opt = (1e-7, 0.999, 1e-4)
    o2, _, _ = oracle.solve(O, P, Y, led, A, r, 2, L, opt, prec="f64", halve_from=0)
    o3, _, _ = oracle.solve(O, P, Y, led, A, r, 3, L, opt, prec="f64", halve_from=0)
    o1, _, _ = oracle.solve(O, P, Y, led, A, r, 1, L, opt, prec="f64", halve_from=0)
    d1 = np.linalg.norm(o2 - o1)
    d2 = np.linalg.norm(o3 - o2)
    beta = float(np.float32(0.999))
    f = lambda k: np.sqrt(1 - beta ** k)  # noqa: E731
    assert abs(d1 / np.linalg.norm(o1 - O.astype(np.complex128)) - 0.5 * f(1) / f(2)) < 1e-3
    assert abs(d2 / d1 - 0.5 * f(2) / f(3)) < 1e-3


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


def _numpy_gradient_with_signs(O, P, Y, led, r, signs):
    """Independent numpy restatement (np.fft, vectorised index maps) of one
    minibatch gradient with the loss stencil's right / down difference signs
    taken from `signs` (bit 0 / bit 1 set: negative)."""
    Nr, Nc = O.shape
    nr, nc = P.shape
    S = np.fft.fft2(O)
    GS = np.zeros_like(S)
    GP = np.zeros_like(P, dtype=np.complex128)
    ii = np.arange(nr)[:, None]
    jj = np.arange(nc)[None, :]
    for l in range(led.shape[1]):
        kl, kt = int(led[0, l]), int(led[1, l])
        rows = (ii + kt - 1 + Nr // 2) % Nr
        cols = (jj + kl - 1 + Nc // 2) % Nc
        sub = S[rows, cols]
        z = np.fft.ifft2(np.fft.ifftshift(sub * P)) * (nr * nc)
        sx = np.where(signs[:, :, l] & 1, -1.0, 1.0)
        sy = np.where(signs[:, :, l] & 2, -1.0, 1.0)
        g = r * r * (np.roll(sx, 1, axis=1) - sx + np.roll(sy, 1, axis=0) - sy)
        mag = np.abs(z)
        u = np.where(mag > 0, g * z / np.where(mag > 0, mag, 1), 0)
        R = np.fft.fftshift(np.fft.fft2(u))
        np.add.at(GS, (np.broadcast_to(rows, (nr, nc)), np.broadcast_to(cols, (nr, nc))), R * np.conj(P) / (Nr * Nc))
        GP += R * np.conj(sub)
    return np.fft.ifft2(GS) * (Nr * Nc), GP


def test_gradient_with_imposed_signs():
    """oracle.gradient_signs: (1) its own decisions imposed reproduce
    oracle.gradient bit for bit, with zero flips; (2) every decision flipped
    negates the gradient exactly (u is linear in the stencil value); (3)
    random decisions imposed match the numpy restatement; (4) the FP32
    oracle's decisions on margin inputs equal the FP64 ones (no flips)."""
    n, r = 16, 2
    O, P, Y, led = margin_problem(n, r)
    GO, GP, loss, _ = oracle.gradient(O, P, Y, led, r)
    g1 = oracle.gradient_signs(O, P, Y, led, r)
    own = g1[3]
    assert own.shape == Y.shape and own.max() <= 3 and 0 < own.mean() < 3
    g2 = oracle.gradient_signs(O, P, Y, led, r, signs=own)
    assert np.array_equal(g2[0], GO) and np.array_equal(g2[1], GP) and g2[4] == 0
    assert abs(g2[2] - loss) <= 1e-12 * loss  # an OpenMP reduction: order may vary
    g3 = oracle.gradient_signs(O, P, Y, led, r, signs=own ^ 3)
    assert np.array_equal(g3[0], -GO) and np.array_equal(g3[1], -GP)
    assert g3[4] == 2 * own.size
    rng = np.random.default_rng(5)
    rnd = np.asfortranarray(rng.integers(0, 4, size=Y.shape).astype(np.uint8))
    g4 = oracle.gradient_signs(O, P, Y, led, r, signs=rnd)
    nO, nP = _numpy_gradient_with_signs(O.astype(np.complex128), P.astype(np.complex128), Y, led, r, rnd)
    assert rel_l2(g4[0], nO) < 1e-12 and rel_l2(g4[1], nP) < 1e-12
    flips = int(np.unpackbits((own ^ rnd)[..., None], axis=-1).sum())
    assert g4[4] == flips
    f32 = oracle.gradient_signs(O, P, Y, led, r, prec="f32")
    assert np.array_equal(f32[3], own)
