"""Shared test helpers: small problems with guaranteed sign margins."""
import numpy as np

import fpm_amd
from fpm_amd import synth


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def small_leds(n, r, grid=3, spiral=True):
    """LED windows of a grid x grid array (the standard optics, fewer LEDs)."""
    led, _, _ = fpm_amd.led_offsets(n, r, led_grid=grid, spiral=spiral)
    return led


def small_leds_oracle(n, r, grid=3, spiral=True):
    import oracle
    led, _, _ = oracle.led_offsets_matlab(n, r, led_grid=grid, spiral=spiral)
    return led


def margin_field(shape, rng, step=0.01, levels=10):
    """Integer-level field k * step with every circular right/down neighbour
    pair different, so every sign decision of the loss stencil has a margin
    of at least `step`."""
    nr, nc = shape
    k = rng.integers(0, levels, size=shape)
    for _ in range(1000):
        bad = (k == np.roll(k, -1, axis=1)) | (k == np.roll(k, -1, axis=0)) | \
              (k == np.roll(k, 1, axis=1)) | (k == np.roll(k, 1, axis=0))
        if not bad.any():
            break
        k[bad] = rng.integers(0, levels, size=int(bad.sum()))
    assert not ((k == np.roll(k, -1, axis=1)) | (k == np.roll(k, -1, axis=0))).any()
    return k * step


def margin_problem(n, r, grid=3, seed=7, spiral=True):
    """(O, P, Y, led) whose residual e = model - Y is a margin field: every
    loss-stencil sign is decided by >= 0.01, far above FP32 rounding."""
    rng = np.random.default_rng(seed)
    N = n * r
    led = small_leds(n, r, grid, spiral)
    _, P = synth.pupil(n)
    amp = synth.object_numpy(N, r, 0.3, seed)
    O = np.asfortranarray(amp.astype(np.complex128))
    model = synth.image_model_numpy(O, P.astype(np.complex128), led, r)
    L = led.shape[1]
    Y = np.empty((n, n, L), dtype=np.float32, order="F")
    for l in range(L):
        Y[:, :, l] = (model[:, :, l] - margin_field((n, n), rng)).astype(np.float32)
    return O.astype(np.complex64), P, Y, led


def texture(n, seed, shift=(0.0, 0.0), width=0.06):
    """Smooth periodic random texture (Gaussian low-pass of white noise,
    frequency width `width` cycles/pixel) translated by shift = (dy, dx)
    through the Fourier shift theorem (exact for any sub-pixel amount)."""
    rng = np.random.default_rng(seed)
    f = np.fft.fft2(rng.standard_normal((n, n)))
    ky = np.fft.fftfreq(n)[:, None]
    kx = np.fft.fftfreq(n)[None, :]
    f *= np.exp(-(ky ** 2 + kx ** 2) / (2 * width ** 2))
    f *= np.exp(-2j * np.pi * (ky * shift[0] + kx * shift[1]))
    return np.real(np.fft.ifft2(f))


def rect_margin_problem(nr, nc, r, L, seed=5):
    """Non-square geometry: object (nr r) x (nc r), pupil nr x nc (an
    elliptical NA disk with a defocus phase), L LED windows around the centre
    (kl indexes columns, kt rows), and images whose residual is a margin
    field (every loss-stencil sign decided by >= 0.01)."""
    rng = np.random.default_rng(seed)
    Nr, Nc = nr * r, nc * r
    spec = np.fft.fft2(rng.random((Nr, Nc)))
    fy = np.fft.fftfreq(Nr)[:, None]
    fx = np.fft.fftfreq(Nc)[None, :]
    spec *= (fy ** 2 + fx ** 2) <= (0.5 / r) ** 2
    amp = np.real(np.fft.ifft2(spec))
    amp = 0.3 + 0.7 * (amp - amp.min()) / (amp.max() - amp.min())
    O = np.asfortranarray((amp * np.exp(1j * rng.uniform(-0.5, 0.5, (Nr, Nc)))).astype(np.complex64))
    yy = (np.arange(nr) - nr // 2)[:, None] / (0.4 * nr)
    xx = (np.arange(nc) - nc // 2)[None, :] / (0.4 * nc)
    rho2 = yy ** 2 + xx ** 2
    A = np.asfortranarray((rho2 <= 1.0).astype(np.float32))
    P = np.asfortranarray((A * np.exp(0.5j * np.sqrt(3) * (2 * rho2 - 1))).astype(np.complex64))
    c_r, c_c = Nr // 2 - nr // 2 + 1, Nc // 2 - nc // 2 + 1
    led = np.asfortranarray(np.stack([c_c + rng.integers(-nc // 2, nc // 2 + 1, L),
                                      c_r + rng.integers(-nr // 2, nr // 2 + 1, L)]).astype(np.int32))
    led[:, 0] = (c_c, c_r)
    model = synth.image_model_numpy(O.astype(np.complex128), P.astype(np.complex128), led, r)
    Y = np.empty((nr, nc, L), dtype=np.float32, order="F")
    for l in range(L):
        Y[:, :, l] = (model[:, :, l] - margin_field((nr, nc), rng)).astype(np.float32)
    return O, P, Y, led, A
