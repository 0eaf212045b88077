"""Gradient-driven solves: the reference's MATLAB-side optimizers on the GPU.

The reference's first two generations drive the gradient-only mex
(fullyfusedFPM, here fpm_fused_gradient) from MATLAB
(codes/main_run_FPM.m:106-160, benchmark/cuFPM_recon_2025_circular.m:
152-183): per minibatch of contiguous LEDs

    [dldw1, dldw2] = fullyfusedFPM(w1, w2, Y(:, :, first:last), ledIdx, pratio);
    w1 = optimizer_w1.step(w1, dldw1);
    w2 = optimizer_w2.step(w2, dldw2);
    w2 = max(min(abs(w2), P0 + ac), P0 - ac) .* sign(w2) .* P0;

The classes below keep the MATLAB constructors and step() of
codes/func_ware/optimizers/optimizer_adam.m, optimizer_RMSprop.m,
optimizer_sgd.m and codes_v2/func_ware/optimizers/optimizer_yogi.m; each
step is one libfpm_amd kernel (fpm_optimizer_step) on device tensors, with
the moments in device memory.  solve_with_optimizers runs the loop above.
"""
from __future__ import annotations

import ctypes
from typing import Optional

from . import _check, lib

OPT_SGD, OPT_ADAM, OPT_YOGI, OPT_RMSPROP = 0, 1, 2, 3

_bound = False


def _olib():
    global _bound
    L = lib()
    if not _bound:
        vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        L.fpm_optimizer_step.argtypes = [i32, vp, vp, vp, vp, i64, dbl, dbl, dbl, i64, vp]
        L.fpm_optimizer_step.restype = ctypes.c_int
        L.fpm_pupil_constraint.argtypes = [vp, vp, i64, dbl, vp]
        L.fpm_pupil_constraint.restype = ctypes.c_int
        _bound = True
    return L


def _stream(t):
    import torch

    return torch.cuda.current_stream(t.device).cuda_stream


def _check_pair(para, grad):
    import torch

    if para.dtype != torch.complex64 or grad.dtype != torch.complex64:
        raise TypeError("parameters and gradients are complex64 tensors")
    if not (para.is_cuda and grad.is_cuda and para.is_contiguous() and grad.is_contiguous()):
        raise ValueError("parameters and gradients are contiguous CUDA tensors")
    if para.shape != grad.shape:
        raise ValueError("gradient shape differs from the parameter's")


class _Moments:
    """MATLAB's mom1 = 0, mom2 = 0 start: device zeros made at the first step."""

    def __init__(self, mom1=0, mom2=0):
        if not (isinstance(mom1, (int, float)) and mom1 == 0 and isinstance(mom2, (int, float)) and mom2 == 0):
            raise ValueError("the moments start at 0 (the reference's callers pass 0, 0)")
        self.mom1 = None
        self.mom2 = None

    def _ensure(self, para):
        import torch

        if self.mom1 is None or self.mom1.shape != para.shape:
            self.mom1 = torch.zeros_like(para)
            self.mom2 = torch.zeros(para.shape, dtype=torch.float32, device=para.device)


class optimizer_adam(_Moments):
    """codes/func_ware/optimizers/optimizer_adam.m:15-38.  step(para, grad,
    iter): iter (1-based) enters the bias correction; without it an internal
    counter is used."""

    name = "adam"

    def __init__(self, mom1=0, mom2=0, beta1=0.9, beta2=0.999, lr=1e-2):
        super().__init__(mom1, mom2)
        self.r1, self.r2, self.lr = float(beta1), float(beta2), float(lr)
        self._iter = 0

    def step(self, para, grad, iter: Optional[int] = None):
        _check_pair(para, grad)
        self._ensure(para)
        self._iter = int(iter) if iter is not None else self._iter + 1
        _check(_olib().fpm_optimizer_step(OPT_ADAM, para.data_ptr(), grad.data_ptr(), self.mom1.data_ptr(),
                                          self.mom2.data_ptr(), para.numel(), self.lr, self.r1, self.r2, self._iter,
                                          _stream(para)))
        return para


class optimizer_yogi(_Moments):
    """codes_v2/func_ware/optimizers/optimizer_yogi.m:15-36."""

    name = "yogi"

    def __init__(self, mom1=0, mom2=0, beta1=0.9, beta2=0.999, lr=1e-2):
        super().__init__(mom1, mom2)
        self.r1, self.r2, self.lr = float(beta1), float(beta2), float(lr)

    def step(self, para, grad):
        _check_pair(para, grad)
        self._ensure(para)
        _check(_olib().fpm_optimizer_step(OPT_YOGI, para.data_ptr(), grad.data_ptr(), self.mom1.data_ptr(),
                                          self.mom2.data_ptr(), para.numel(), self.lr, self.r1, self.r2, 0,
                                          _stream(para)))
        return para


class optimizer_RMSprop:
    """codes/func_ware/optimizers/optimizer_RMSprop.m:18-46 as every caller
    constructs it, optimizer_RMSprop(0, 0, beta1, 0, false, lr): squave =
    beta1 + (1 - beta1)|g|^2 (no running average, as written), step
    g / (sqrt(squave) + 1e-5).  Other constructions are not supported."""

    name = "RMSprop"

    def __init__(self, squave=0, buffer=0, beta1=0.9, mum=0, centered=False, lr=1e-2):
        if squave != 0 or buffer != 0 or mum != 0 or centered:
            raise NotImplementedError("supported: optimizer_RMSprop(0, 0, beta1, 0, false, lr), as the callers use it")
        self.r1, self.lr = float(beta1), float(lr)

    def step(self, para, grad):
        _check_pair(para, grad)
        _check(_olib().fpm_optimizer_step(OPT_RMSPROP, para.data_ptr(), grad.data_ptr(), None, None, para.numel(),
                                          self.lr, self.r1, 0.0, 0, _stream(para)))
        return para


class optimizer_sgd:
    """codes/func_ware/optimizers/optimizer_sgd.m:11-21: para - lr grad (the
    reference subtracts an undefined `update`; the gradient step is meant)."""

    name = "sgd"

    def __init__(self, lr=1e-2):
        self.lr = float(lr)

    def step(self, para, grad):
        _check_pair(para, grad)
        _check(_olib().fpm_optimizer_step(OPT_SGD, para.data_ptr(), grad.data_ptr(), None, None, para.numel(),
                                          self.lr, 0.0, 0.0, 0, _stream(para)))
        return para


def pupil_constraint(P, A, ac: float):
    """In place: P = max(min(|P|, A + ac), A - ac) . sign(P) . A
    (codes/main_run_FPM.m:153-156).  P complex64, A float32, CUDA tensors."""
    import torch

    if P.dtype != torch.complex64 or A.dtype != torch.float32 or P.shape != A.shape:
        raise ValueError("P complex64 and A float32 of the same shape")
    if not (P.is_cuda and A.is_cuda and P.is_contiguous() and A.is_contiguous()):
        raise ValueError("contiguous CUDA tensors")
    _check(_olib().fpm_pupil_constraint(P.data_ptr(), A.data_ptr(), P.numel(), float(ac), _stream(P)))
    return P


def solve_with_optimizers(O0, P0, Y, led, A, pratio: int, epochs: int, batch: int, opt_O, opt_P, ac: float = 0.1,
                          *, stitch: Optional[int] = None, engine: int = 0, loss: bool = False):
    """The gradient-driven loop of codes/main_run_FPM.m:106-160 on device
    tensors in the solve's layout (row-major buffers of the column-major
    arrays, as the rest of the package): O0 (N, N) complex64, P0 (n, n)
    complex64, Y (L, n, n) float32 amplitudes (prepared once over the whole
    stack, as the reference does before its loop), led (L, 2) int32 = the
    2 x L column-major {kl, kt}, A (n, n) float32.  Minibatches are the
    contiguous LED chunks first..first+batch-1; per chunk one fused gradient,
    one step of opt_O on O and of opt_P on P, then the pupil constraint with
    amplitude bound ac.  Returns (O, P, per-step losses or None)."""
    import torch

    from . import STITCH_OWNER, gradient_device

    if Y.dtype != torch.float32:
        raise TypeError("Y: float32 amplitudes (prepare a uint16 stack once with fpm_prepare_amplitude)")
    for t in (O0, P0, Y, led, A):
        if not (t.is_cuda and t.is_contiguous()):
            raise ValueError("device tensors, contiguous")
    N, n, L = int(O0.shape[0]), int(P0.shape[0]), int(Y.shape[0])
    if led.shape != (L, 2):
        raise ValueError("led: (L, 2) int32")
    O, P = O0.clone(), P0.clone()
    GO, GP = torch.empty_like(O), torch.empty_like(P)
    plane = n * n
    losses = [] if loss else None
    for _ in range(int(epochs)):
        first = 0
        while first < L:
            b = min(int(batch), L - first)
            lv = gradient_device(O.data_ptr(), P.data_ptr(), Y.data_ptr() + 4 * plane * first, False,
                                 led.data_ptr() + 8 * first, (N, N), (n, n), b, pratio, GO.data_ptr(), GP.data_ptr(),
                                 stitch=STITCH_OWNER if stitch is None else stitch,
                                 stream=torch.cuda.current_stream(O.device).cuda_stream, engine=engine)
            if losses is not None:
                losses.append(lv)
            opt_O.step(O, GO)
            opt_P.step(P, GP)
            pupil_constraint(P, A, ac)
            first += b
    torch.cuda.synchronize(O.device)
    return O, P, losses
