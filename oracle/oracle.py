"""oracle.py -- numpy face of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker of libfpm_amd.  Wraps oracle/build/liboracle.so (C
restatement of THUHoloLab/Fused-Fourier-Ptycho's cuFPM_v3 path, see
fpm_oracle.c) and adds a numpy restatement of the LED geometry of
codes_v2/func_ware/init_environment_rgb.m:33-134 that mirrors MATLAB's own
operations (sin/cos + round in the spiral, meshgrid, find(min)), written
independently of the product's C++ version so the two check each other.

Arrays use MATLAB indexing (row, col[, page]) in Fortran order.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "build", "liboracle.so")
_lib = None


def build() -> str:
    """Compile the oracle (gcc, OpenMP) into oracle/build/."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _SO


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        vp, i, d, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_float
        for s in ("f64", "f32"):
            getattr(L, f"oracle_fft2_{s}").argtypes = [vp, i, i, i, i]
            getattr(L, f"oracle_gradient_{s}").argtypes = [vp, i, i, vp, i, i, vp, i, vp, i, vp, vp,
                                                           ctypes.POINTER(d), ctypes.POINTER(d)]
            getattr(L, f"oracle_gradient_signs_{s}").argtypes = [vp, i, i, vp, i, i, vp, i, vp, i, vp, vp,
                                                                 ctypes.POINTER(d), ctypes.POINTER(d), vp, vp,
                                                                 ctypes.POINTER(ctypes.c_long)]
            getattr(L, f"oracle_loss_{s}").argtypes = [vp, i, i, vp, i, i, vp, i, vp, i]
            getattr(L, f"oracle_loss_{s}").restype = d
            getattr(L, f"oracle_solve_{s}").argtypes = [vp, i, i, vp, i, i, vp, i, vp, vp, i, i, i, f, f, f, f,
                                                        vp, vp, vp, i, i]
        L.oracle_prepare_amplitude.argtypes = [vp, ctypes.c_long, vp]
        L.oracle_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _ct(prec):
    return np.complex128 if prec == "f64" else np.complex64


def _fa(a, dt):
    return np.asfortranarray(np.asarray(a, dtype=dt))


def num_threads() -> int:
    return int(lib().oracle_num_threads())


def fft2(x, inverse=False, prec="f64"):
    """Unnormalised 2D transform of each (rows, cols) page."""
    a = _fa(x, _ct(prec)).copy(order="F")
    rows, cols = a.shape[:2]
    batch = int(np.prod(a.shape[2:])) if a.ndim > 2 else 1
    getattr(lib(), f"oracle_fft2_{prec}")(a.ctypes.data, rows, cols, batch, 1 if inverse else 0)
    return a


def gradient(O, P, Y, led, pratio, prec="f64"):
    """(G_O, G_P, loss, margin) of one minibatch made of all pages of Y."""
    O = _fa(O, _ct(prec))
    P = _fa(P, _ct(prec))
    Y = _fa(Y, np.float32)
    if Y.ndim == 2:
        Y = Y[:, :, None].copy(order="F")
    led = _fa(led, np.int32)
    GO = np.zeros_like(O, order="F")
    GP = np.zeros_like(P, order="F")
    loss, margin = ctypes.c_double(0), ctypes.c_double(0)
    getattr(lib(), f"oracle_gradient_{prec}")(O.ctypes.data, O.shape[0], O.shape[1], P.ctypes.data, P.shape[0],
                                              P.shape[1], Y.ctypes.data, Y.shape[2], led.ctypes.data, int(pratio),
                                              GO.ctypes.data, GP.ctypes.data, ctypes.byref(loss), ctypes.byref(margin))
    return GO, GP, loss.value, margin.value


def gradient_signs(O, P, Y, led, pratio, signs=None, prec="f64"):
    """(G_O, G_P, loss, own_signs, flips) of one minibatch made of all pages
    of Y with the loss stencil's sign decisions taken from `signs` (uint8 n_r
    x n_c x L, bit 0: sign bit of e(i, j+1) - e(i, j), bit 1: of e(i+1, j) -
    e(i, j)) instead of its own; own_signs are the decisions it would have
    taken and flips the number of bits in which they differ from `signs`."""
    O = _fa(O, _ct(prec))
    P = _fa(P, _ct(prec))
    Y = _fa(Y, np.float32)
    if Y.ndim == 2:
        Y = Y[:, :, None].copy(order="F")
    led = _fa(led, np.int32)
    GO = np.zeros_like(O, order="F")
    GP = np.zeros_like(P, order="F")
    own = np.zeros(Y.shape, dtype=np.uint8, order="F")
    sg = None if signs is None else _fa(signs, np.uint8)
    if sg is not None and sg.shape != Y.shape:
        raise ValueError(f"signs must be {Y.shape}, got {sg.shape}")
    loss, margin, flips = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_long(0)
    getattr(lib(), f"oracle_gradient_signs_{prec}")(
        O.ctypes.data, O.shape[0], O.shape[1], P.ctypes.data, P.shape[0], P.shape[1], Y.ctypes.data, Y.shape[2],
        led.ctypes.data, int(pratio), GO.ctypes.data, GP.ctypes.data, ctypes.byref(loss), ctypes.byref(margin),
        None if sg is None else sg.ctypes.data, own.ctypes.data, ctypes.byref(flips))
    return GO, GP, loss.value, own, int(flips.value)


def loss(O, P, Y, led, pratio, prec="f64"):
    O = _fa(O, _ct(prec))
    P = _fa(P, _ct(prec))
    Y = _fa(Y, np.float32)
    if Y.ndim == 2:
        Y = Y[:, :, None].copy(order="F")
    led = _fa(led, np.int32)
    return getattr(lib(), f"oracle_loss_{prec}")(O.ctypes.data, O.shape[0], O.shape[1], P.ctypes.data, P.shape[0],
                                                 P.shape[1], Y.ctypes.data, Y.shape[2], led.ctypes.data, int(pratio))


def solve(O0, P0, Y, led, A, pratio, epochs, batch, opt, prec="f32", clamp=0.2, max_steps=0, halve_from=13):
    """The v3 solve restated on the CPU: (O, P, loss_per_epoch); lr halves
    after every epoch with index >= halve_from (13: the reference)."""
    O0 = _fa(O0, _ct(prec))
    P0 = _fa(P0, _ct(prec))
    Y = _fa(Y, np.float32)
    led = _fa(led, np.int32)
    A = _fa(A, np.float32)
    O = np.empty_like(O0, order="F")
    P = np.empty_like(P0, order="F")
    lpe = np.zeros(max(int(epochs), 1), dtype=np.float64)
    getattr(lib(), f"oracle_solve_{prec}")(O0.ctypes.data, O0.shape[0], O0.shape[1], P0.ctypes.data, P0.shape[0],
                                           P0.shape[1], Y.ctypes.data, Y.shape[2], led.ctypes.data, A.ctypes.data,
                                           int(pratio), int(epochs), int(batch), float(opt[0]), float(opt[1]),
                                           float(opt[2]), float(clamp), O.ctypes.data, P.ctypes.data,
                                           lpe.ctypes.data, int(max_steps), int(halve_from))
    return O, P, lpe[: int(epochs)]


def prepare_amplitude(raw):
    r = _fa(raw, np.uint16)
    out = np.empty(r.shape, dtype=np.float32, order="F")
    lib().oracle_prepare_amplitude(r.ctypes.data, r.size, out.ctypes.data)
    return out


def led_offsets_matlab(n, pratio, wavelength_um=0.532, pixel_um=1.625, led_pitch_um=4000.0,
                       led_height_um=90000.0, led_grid=19, spiral=True, objdx=0.0, objdy=0.0):
    """numpy restatement of init_environment_rgb.m:33-134 + main_run_FPM.m:23,27-31,53-55.

    Mirrors MATLAB step by step, including sin/cos(pi/2*d) + round in the
    spiral and find(x == min(x)); returns (ledIdx int32 2 x L, grid k, min
    relative argmin margin)."""
    lam = wavelength_um
    k_lamuda = 2 * np.pi / lam
    N = M = n
    NN = MM = n * pratio
    D_pixel = pixel_um
    Fxx1 = np.arange(-(NN // 2), NN // 2, dtype=np.float64) / (N * D_pixel) * (2 * np.pi)
    Fyy1 = np.arange(-(MM // 2), MM // 2, dtype=np.float64) / (M * D_pixel) * (2 * np.pi)
    g = led_grid
    lit_cen = (g - 1) / 2
    vled = np.arange(g) - lit_cen
    hled = np.arange(g) - lit_cen
    hhled, vvled = np.meshgrid(hled, vled)
    # init_environment_rgb.m:83-84 (objdx, objdy: the field area, :26-29)
    den = np.sqrt((-vvled * led_pitch_um + objdx) ** 2 + (-hhled * led_pitch_um + objdy) ** 2 + led_height_um ** 2)
    v = (-vvled * led_pitch_um + objdx) / den
    u = -(-hhled * led_pitch_um + objdy) / den
    u, v = v, u  # init_environment_rgb.m:86-88
    pos = np.zeros((g, g, 2), dtype=np.int64)
    margin = np.inf
    step = Fxx1[1] - Fxx1[0]
    for i in range(g):
        for j in range(g):
            d1 = np.abs(Fxx1 - k_lamuda * u[i, j])
            d2 = np.abs(Fyy1 - k_lamuda * v[i, j])
            w1 = np.flatnonzero(d1 == d1.min())
            w2 = np.flatnonzero(d2 == d2.min())
            assert w1.size == 1 and w2.size == 1, "find(min) returned a tie"
            pos[i, j, 0] = w1[0] + 1
            pos[i, j, 1] = w2[0] + 1
            s1 = np.sort(d1)[:2]
            s2 = np.sort(d2)[:2]
            margin = min(margin, (s1[1] - s1[0]) / step, (s2[1] - s2[0]) / step)
    # spiral order (init_environment_rgb.m:110-134), MATLAB 1-based
    if spiral:
        seen = np.zeros((g, g), dtype=bool)
        ii, jj = (g + 1) // 2, (g + 1) // 2
        oi, oj = [ii], [jj]
        seen[ii - 1, jj - 1] = True
        direction = 0
        while not seen.all():
            d2_ = direction + 1
            ii2 = int(np.round(ii + np.sin(np.pi / 2 * d2_)))
            jj2 = int(np.round(jj + np.cos(np.pi / 2 * d2_)))
            if not seen[ii2 - 1, jj2 - 1]:
                direction = d2_
            ii = int(np.round(ii + np.sin(np.pi / 2 * direction)))
            jj = int(np.round(jj + np.cos(np.pi / 2 * direction)))
            oi.append(ii)
            oj.append(jj)
            seen[ii - 1, jj - 1] = True
    else:
        oi = [i + 1 for i in range(g) for _ in range(g)]
        oj = [j + 1 for _ in range(g) for j in range(g)]
    uo = np.array([pos[i - 1, j - 1, 0] for i, j in zip(oi, oj)])
    vo = np.array([pos[i - 1, j - 1, 1] for i, j in zip(oi, oj)])
    kt = -M // 2 + vo
    kl = -N // 2 + uo
    led = np.asfortranarray(np.stack([kl, kt]).astype(np.int32))
    k = np.array([j + (i - 1) * g for i, j in zip(oi, oj)], dtype=np.int32)
    return led, k, margin
