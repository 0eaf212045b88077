"""fpm_amd -- Python mirror of the cuFPM_pure_v3 mex interface over libfpm_amd.so.

The reference is called from MATLAB as (codes_v2/main_run_FPM.m:98-106,
cuFPM_v3/cuFPM_pure_v3.cu:44-228 of THUHoloLab/Fused-Fourier-Ptycho)::

    [O, P] = cuFPM_pure_v3(O0, P0, Y, ledIdx, A, pratio, numEpochs, batch, [lr beta eps])

:func:`cuFPM_pure_v3` takes the same arguments with MATLAB's meaning: arrays are
indexed (row, col[, page]) and stored column-major (numpy Fortran order), ``O0``
is the spatial-domain object, ``ledIdx`` is int32 2 x L of {kl; kt} 1-based
window starts in the centred spectrum.  Numeric scalars are truncated to int as
the mex does.  The inputs are never modified; new arrays are returned.

Device-resident data (torch tensors on a GPU, or raw device pointers) go
through :func:`solve_device`, which is what the benchmark times.

All compute runs in the HIP library; there is no CPU fallback: if the library
is missing or no GPU is present, calls raise :class:`FPMError`.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence, Tuple

import numpy as np

__all__ = [
    "FPMError", "lib", "library_path", "cuFPM_pure_v3", "solve_FPM", "fused_gradient", "led_offsets",
    "prepare_amplitude", "forward_intensity", "fft2", "solve_device", "gradient_device",
    "STITCH_ATOMIC", "STITCH_OWNER", "Y_F32_AMPLITUDE", "Y_U16_RAW", "EXPORTED_SYMBOLS",
]

STITCH_ATOMIC = 0
STITCH_OWNER = 1
FFT_F32 = 0        # FP32 butterflies (default)
FFT_SPLIT_F16 = 1  # radix-16 DFTs as split hi+lo FP16 MFMA products, FP32 accumulation
Y_F32_AMPLITUDE = 0
Y_U16_RAW = 1

# every symbol include/fpm_amd.h declares
EXPORTED_SYMBOLS = (
    "fpm_opts_init", "fpm_pure_v3", "fpm_fused_gradient", "fpm_led_offsets", "fpm_prepare_amplitude",
    "fpm_forward_intensity", "fpm_fft2", "fpm_fft2_engine", "fpm_last_error", "fpm_abi_version",
    "fpm_cnn_create", "fpm_cnn_destroy", "fpm_cnn_num_params", "fpm_cnn_get_params", "fpm_cnn_set_params",
    "fpm_cnn_forward", "fpm_cnn_gradient", "fpm_cnn_train_step", "fpm_led_offsets_at", "fpm_tiles_solve",
    "fpm_parallax_shifts", "fpm_optimizer_step", "fpm_pupil_constraint", "fpm_fused_gradient_noalloc",
)

_HERE = os.path.dirname(os.path.abspath(__file__))
# FPM_AMD_LIBRARY selects another build of the same library (e.g. a variant
# compiled with different tile constants); default: the in-tree build
library_path = os.environ.get("FPM_AMD_LIBRARY") or os.path.join(_HERE, "libfpm_amd.so")


class FPMError(RuntimeError):
    """Raised when a libfpm_amd entry point returns a non-zero status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


STAGE_COUNT = 13


class StageTimes(ctypes.Structure):
    """fpm_stage_times: per-stage device milliseconds, executions, launches."""
    _fields_ = [
        ("ms", ctypes.c_double * STAGE_COUNT),
        ("count", ctypes.c_int64 * STAGE_COUNT),
        ("launches", ctypes.c_int32 * STAGE_COUNT),
    ]


class _Opts(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_int32),
        ("stitch", ctypes.c_int32),
        ("pointers_are_device", ctypes.c_int32),
        ("verbose", ctypes.c_int32),
        ("stream", ctypes.c_void_p),
        ("loss_per_epoch", ctypes.POINTER(ctypes.c_double)),
        ("pupil_clamp", ctypes.c_float),
        ("stage_times", ctypes.POINTER(StageTimes)),
        ("graph", ctypes.c_int32),
        ("fft_engine", ctypes.c_int32),
        ("lr_halve_from", ctypes.c_int32),
        ("tile_launch", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("stop_loss", ctypes.c_double),
        ("epochs_run", ctypes.POINTER(ctypes.c_int32)),
        ("stencil_signs", ctypes.c_void_p),
    ]


_lib = None


def lib() -> ctypes.CDLL:
    """Load libfpm_amd.so (built in-tree by ``make -C fused-fourier-ptycho_amd``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(library_path):
        raise FPMError(-5, f"{library_path} is not built (run make -C fused-fourier-ptycho_amd)")
    L = ctypes.CDLL(library_path)
    vp, i64, i32, dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_double
    L.fpm_opts_init.argtypes = [ctypes.POINTER(_Opts)]
    L.fpm_opts_init.restype = None
    L.fpm_pure_v3.argtypes = [vp, i64, i64, vp, i64, i64, vp, i32, i64, vp, vp, dbl, dbl, dbl,
                              ctypes.POINTER(ctypes.c_double), vp, vp, ctypes.POINTER(_Opts)]
    L.fpm_pure_v3.restype = ctypes.c_int
    L.fpm_fused_gradient.argtypes = [vp, i64, i64, vp, i64, i64, vp, i32, i64, vp, dbl, vp, vp,
                                     ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_Opts)]
    L.fpm_fused_gradient.restype = ctypes.c_int
    L.fpm_led_offsets.argtypes = [dbl, dbl, dbl, dbl, dbl, i32, i64, i32, i32, vp, vp, vp]
    L.fpm_led_offsets.restype = ctypes.c_int
    L.fpm_prepare_amplitude.argtypes = [vp, i64, vp, i32, vp, ctypes.POINTER(ctypes.c_uint16),
                                        ctypes.POINTER(ctypes.c_uint16)]
    L.fpm_prepare_amplitude.restype = ctypes.c_int
    L.fpm_forward_intensity.argtypes = [vp, i64, i64, vp, i64, i64, i64, vp, dbl, i32, vp, i32, vp]
    L.fpm_forward_intensity.restype = ctypes.c_int
    L.fpm_fft2.argtypes = [vp, vp, i64, i64, i64, i32, i32, vp]
    L.fpm_fft2.restype = ctypes.c_int
    L.fpm_fft2_engine.argtypes = [vp, vp, i64, i64, i64, i32, i32, i32, vp]
    L.fpm_fft2_engine.restype = ctypes.c_int
    L.fpm_fused_gradient_noalloc.argtypes = [vp, i64, i64, vp, i64, i64, vp, i64, vp, dbl, vp, vp, vp, vp]
    L.fpm_fused_gradient_noalloc.restype = ctypes.c_int
    L.fpm_last_error.argtypes = []
    L.fpm_last_error.restype = ctypes.c_char_p
    L.fpm_abi_version.argtypes = []
    L.fpm_abi_version.restype = ctypes.c_int32
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != 0:
        raise FPMError(rc, lib().fpm_last_error().decode(errors="replace"))


def _opts(stitch=STITCH_OWNER, device=False, stream=None, loss_buf=None, clamp=0.2, verbose=False,
          graph=False, engine=FFT_F32, lr_halve_from=0, stop_loss=0.0, epochs_run=None,
          tile_launch=0) -> _Opts:
    o = _Opts()
    lib().fpm_opts_init(ctypes.byref(o))
    o.stitch = int(stitch)
    o.pointers_are_device = 1 if device else 0
    o.verbose = 1 if verbose else 0
    o.stream = stream
    if loss_buf is not None:
        o.loss_per_epoch = loss_buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    o.pupil_clamp = float(clamp)
    o.graph = 1 if graph else 0
    o.fft_engine = int(engine)
    o.lr_halve_from = int(lr_halve_from)
    o.stop_loss = float(stop_loss)
    o.tile_launch = int(tile_launch)
    if epochs_run is not None:
        o.epochs_run = ctypes.pointer(epochs_run)
    return o


def _fc(a, dtype) -> np.ndarray:
    """Fortran-ordered (column-major) copy/view of `a` with dtype."""
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def cuFPM_pure_v3(O0, P0, Y, ledIdx, A, pratio, numEpochs, batch, opt: Sequence[float],
                  *, stitch: int = STITCH_OWNER, loss: bool = False, clamp: float = 0.2,
                  verbose: bool = False, graph: bool = False, engine: int = FFT_F32):
    """The drop-in solve (host arrays).  Returns (O, P), or (O, P, loss_per_epoch).

    ``graph=True`` captures each epoch once as a hipGraph and replays it;
    ``engine=FFT_SPLIT_F16`` runs the radix-16 DFT stages on the matrix cores.

    ``Y`` may be float32 amplitudes (exactly what v3 takes) or uint16 raw
    counts (normalised on the device as codes_v2/main_run_FPM.m:43-45 does).
    """
    O0 = _fc(O0, np.complex64)
    P0 = _fc(P0, np.complex64)
    Yarr = np.asarray(Y)
    ytype = Y_U16_RAW if Yarr.dtype == np.uint16 else Y_F32_AMPLITUDE
    Yarr = _fc(Yarr, np.uint16 if ytype == Y_U16_RAW else np.float32)
    led = _fc(ledIdx, np.int32)
    A = _fc(A, np.float32)
    if O0.ndim != 2 or P0.ndim != 2:
        raise FPMError(-1, "O0 and P0 must be 2-D")
    nr, nc = P0.shape
    if Yarr.ndim == 2:
        Yarr = Yarr.reshape(nr, nc, 1, order="F")
    if Yarr.shape[:2] != (nr, nc) or A.shape != (nr, nc):
        raise FPMError(-1, "Y pages and A must match the pupil shape")
    L = Yarr.shape[2]
    if led.shape != (2, L):
        raise FPMError(-1, f"ledIdx must be int32 2 x L (L = {L}), got {led.shape}")
    ne = int(numEpochs)
    O = np.empty_like(O0, order="F")
    P = np.empty_like(P0, order="F")
    optv = (ctypes.c_double * 3)(*[float(v) for v in opt[:3]])
    lossb = np.zeros(max(ne, 1), dtype=np.float64) if loss else None
    o = _opts(stitch=stitch, loss_buf=lossb, clamp=clamp, verbose=verbose, graph=graph, engine=engine)
    _check(lib().fpm_pure_v3(_ptr(O0), O0.shape[0], O0.shape[1], _ptr(P0), nr, nc, _ptr(Yarr), ytype, L,
                             _ptr(led), _ptr(A), float(pratio), float(numEpochs), float(batch), optv,
                             _ptr(O), _ptr(P), ctypes.byref(o)))
    if loss:
        return O, P, lossb[:ne]
    return O, P


def solve_FPM(target, pupil, obsY, ledidx, circ, *, pratio: float = 8, epoch: float = 20, batchSize: float = 26,
              learningRate: float = 0.0005, beta: float = 0.999, eps: float = 1e-4, **kw):
    """The MATLAB caller of the mex, solve_FPM(parms) of
    codes_v2/main_run_FPM.m:77-107, with its name-value defaults (pratio 8,
    epoch 20, batchSize 26, learningRate 5e-4, beta 0.999, eps 1e-4): returns
    (w1, w2) = cuFPM_pure_v3(target, pupil, obsY, ledidx, circ, pratio, epoch,
    batchSize, [learningRate beta eps]); further keywords go to
    :func:`cuFPM_pure_v3`."""
    circ = np.asarray(circ, dtype=np.float32)
    return cuFPM_pure_v3(target, pupil, obsY, ledidx, circ, pratio, epoch, batchSize,
                         (learningRate, beta, eps), **kw)


def fused_gradient(O, P, Y, ledIdx, pratio, *, stitch: int = STITCH_OWNER, engine: int = FFT_F32,
                   signs: bool = False):
    """One fused gradient over all images: (G_O, G_P, loss), or (G_O, G_P,
    loss, signs) with ``signs=True``: the loss stencil's sign decisions, uint8
    n_r x n_c x L (bit 0: sign bit of e(i, j+1) - e(i, j), bit 1: of
    e(i+1, j) - e(i, j); fpm_opts.stencil_signs).

    The gradient-only call of benchmark/cuFPM/fullyfusedFPM.cu:21-162 with the
    v3 loss: G_O = r^2 grad_O L and G_P = n^2 r^4 grad_P L (N = n r)."""
    O = _fc(O, np.complex64)
    P = _fc(P, np.complex64)
    Yarr = np.asarray(Y)
    ytype = Y_U16_RAW if Yarr.dtype == np.uint16 else Y_F32_AMPLITUDE
    Yarr = _fc(Yarr, np.uint16 if ytype == Y_U16_RAW else np.float32)
    nr, nc = P.shape
    if Yarr.ndim == 2:
        Yarr = Yarr.reshape(nr, nc, 1, order="F")
    L = Yarr.shape[2]
    led = _fc(ledIdx, np.int32)
    if led.shape != (2, L):
        raise FPMError(-1, f"ledIdx must be int32 2 x L (L = {L})")
    GO = np.empty_like(O, order="F")
    GP = np.empty_like(P, order="F")
    lossv = ctypes.c_double(0.0)
    o = _opts(stitch=stitch, engine=engine)
    sg = np.zeros((nr, nc, L), dtype=np.uint8, order="F") if signs else None
    if signs:
        o.stencil_signs = sg.ctypes.data
    _check(lib().fpm_fused_gradient(_ptr(O), O.shape[0], O.shape[1], _ptr(P), nr, nc, _ptr(Yarr), ytype, L,
                                    _ptr(led), float(pratio), _ptr(GO), _ptr(GP), ctypes.byref(lossv),
                                    ctypes.byref(o)))
    if signs:
        return GO, GP, lossv.value, sg
    return GO, GP, lossv.value


def led_offsets(n: int, pratio: int, *, wavelength_um=0.532, na=0.1, pixel_um=1.625, led_pitch_um=4000.0,
                led_height_um=90000.0, led_grid=19, spiral=True):
    """LED window offsets {kl; kt} (int32 2 x L), grid numbers and illumination NA
    (codes_v2/func_ware/init_environment_rgb.m:33-134, main_run_FPM.m:23,53-55)."""
    L = led_grid * led_grid
    led = np.zeros((2, L), dtype=np.int32, order="F")
    gi = np.zeros(L, dtype=np.int32)
    nai = np.zeros(L, dtype=np.float64)
    _check(lib().fpm_led_offsets(float(wavelength_um), float(na), float(pixel_um), float(led_pitch_um),
                                 float(led_height_um), int(led_grid), int(n), int(pratio), 1 if spiral else 0,
                                 _ptr(led), _ptr(gi), _ptr(nai)))
    return led, gi, nai


def prepare_amplitude(raw) -> Tuple[np.ndarray, int, int]:
    """uint16 stack -> float amplitudes sqrt((I-min)/(max-min)) (main_run_FPM.m:43-45)."""
    r = _fc(raw, np.uint16)
    out = np.empty(r.shape, dtype=np.float32, order="F")
    lo, hi = ctypes.c_uint16(0), ctypes.c_uint16(0)
    _check(lib().fpm_prepare_amplitude(_ptr(r), r.size, _ptr(out), 0, None, ctypes.byref(lo), ctypes.byref(hi)))
    return out, lo.value, hi.value


def forward_intensity(O, P, ledIdx, pratio, *, squared=False) -> np.ndarray:
    """Image model |IDFT(ifftshift(crop(DFT O) P))|/(n^2 r^2) per LED (n x n x L float)."""
    O = _fc(O, np.complex64)
    P = _fc(P, np.complex64)
    led = _fc(ledIdx, np.int32)
    L = led.shape[1]
    nr, nc = P.shape
    out = np.empty((nr, nc, L), dtype=np.float32, order="F")
    _check(lib().fpm_forward_intensity(_ptr(O), O.shape[0], O.shape[1], _ptr(P), nr, nc, L, _ptr(led),
                                       float(pratio), 1 if squared else 0, _ptr(out), 0, None))
    return out


def fft2(x, inverse: bool = False, engine: int = FFT_F32) -> np.ndarray:
    """Unnormalised 2D C2C transform of each (rows, cols) page of a Fortran array."""
    a = _fc(x, np.complex64)
    if a.ndim == 2:
        shape, batch = a.shape, 1
    else:
        shape, batch = a.shape[:2], int(np.prod(a.shape[2:]))
    out = np.empty_like(a, order="F")
    _check(lib().fpm_fft2_engine(_ptr(a), _ptr(out), shape[0], shape[1], batch, 1 if inverse else -1, int(engine),
                                 0, None))
    return out


# ----------------------------------------------------------------- device API
def solve_device(O0_ptr: int, P0_ptr: int, Y_ptr: int, y_u16: bool, led_ptr: int, A_ptr: int,
                 N: Tuple[int, int], n: Tuple[int, int], L: int, pratio: float, epochs: float, batch: float,
                 opt: Sequence[float], O_out_ptr: int, P_out_ptr: int, *, stitch: int = STITCH_OWNER,
                 stream: Optional[int] = None, loss: bool = False, clamp: float = 0.2, graph: bool = False,
                 engine: int = FFT_F32, lr_halve_from: int = 0, stop_loss: float = 0.0):
    """fpm_pure_v3 on device pointers (column-major buffers already in HBM).
    Returns the per-epoch loss array of the epochs run if ``loss`` else
    None; with stop_loss > 0 the solve stops after the first epoch whose
    loss is <= stop_loss."""
    optv = (ctypes.c_double * 3)(*[float(v) for v in opt[:3]])
    ne = int(epochs)
    lossb = np.zeros(max(ne, 1), dtype=np.float64) if loss else None
    ran = ctypes.c_int32(0)
    o = _opts(stitch=stitch, device=True, stream=stream, loss_buf=lossb, clamp=clamp, graph=graph, engine=engine,
              lr_halve_from=lr_halve_from, stop_loss=stop_loss, epochs_run=ran)
    _check(lib().fpm_pure_v3(O0_ptr, N[0], N[1], P0_ptr, n[0], n[1], Y_ptr, Y_U16_RAW if y_u16 else Y_F32_AMPLITUDE,
                             L, led_ptr, A_ptr, float(pratio), float(epochs), float(batch), optv, O_out_ptr,
                             P_out_ptr, ctypes.byref(o)))
    return lossb[:ran.value] if loss else None


def gradient_device(O_ptr: int, P_ptr: int, Y_ptr: int, y_u16: bool, led_ptr: int, N, n, L: int, pratio: float,
                    GO_ptr: int, GP_ptr: int, *, stitch: int = STITCH_OWNER, stream: Optional[int] = None,
                    engine: int = FFT_F32, signs_ptr: Optional[int] = None) -> float:
    """fpm_fused_gradient on device pointers; returns the loss.  signs_ptr
    (optional device pointer, n_r n_c L bytes) receives the loss stencil's
    sign decisions (fpm_opts.stencil_signs)."""
    lossv = ctypes.c_double(0.0)
    o = _opts(stitch=stitch, device=True, stream=stream, engine=engine)
    if signs_ptr:
        o.stencil_signs = signs_ptr
    _check(lib().fpm_fused_gradient(O_ptr, N[0], N[1], P_ptr, n[0], n[1], Y_ptr, Y_U16_RAW if y_u16 else Y_F32_AMPLITUDE,
                                    L, led_ptr, float(pratio), GO_ptr, GP_ptr, ctypes.byref(lossv), ctypes.byref(o)))
    return lossv.value
