"""optim_oracle.py -- numpy restatement of the reference's MATLAB-side
optimizers and pupil constraint (TEST INFRASTRUCTURE ONLY: loaded by tests/
as the checker of fpm_optimizer_step / fpm_pupil_constraint and of
fpm_amd.optimizers.solve_with_optimizers).

Follows, in single precision as MATLAB evaluates them on single gpuArrays:
  optimizer_adam     codes/func_ware/optimizers/optimizer_adam.m:25-38
  optimizer_yogi     codes_v2/func_ware/optimizers/optimizer_yogi.m:25-36
  optimizer_RMSprop  codes/func_ware/optimizers/optimizer_RMSprop.m:28-46
                     (centered false, mum 0: squave = r1 + (1 - r1)|g|^2)
  optimizer_sgd      codes/func_ware/optimizers/optimizer_sgd.m:16-21
                     (para - lr grad; the file subtracts an undefined `update`)
  pupil constraint   codes/main_run_FPM.m:153-156
  the loop           codes/main_run_FPM.m:106-160
Scalar rules: a double scalar meeting a single array is rounded to single;
(1 - beta) and Adam's bias are formed in double and rounded once.  |g| is
sqrt(x*x + y*y) in single, every operation rounded on its own (no fused
multiply-add).
"""
from __future__ import annotations

import numpy as np

f32 = np.float32


def _abs2(g):
    ag = np.sqrt(g.real * g.real + g.imag * g.imag)
    return ag * ag


def _sign(x):
    return np.sign(x).astype(f32)  # MATLAB sign(0) = 0


def _cdiv(z, d):
    return (z.real / d + 1j * (z.imag / d)).astype(np.complex64)


def _cscale(s, z):
    return (s * z.real + 1j * (s * z.imag)).astype(np.complex64)


def _csub(a, b):
    return (a.real - b.real + 1j * (a.imag - b.imag)).astype(np.complex64)


def step(kind, para, grad, mom1=None, mom2=None, lr=1e-2, beta1=0.9, beta2=0.999, it=1):
    """One step; returns (para, mom1, mom2) as new arrays (complex64, float32).
    kind: 'adam', 'yogi', 'rmsprop', 'sgd'."""
    para = np.asarray(para, np.complex64)
    grad = np.asarray(grad, np.complex64)
    lr32, b1, b2 = f32(lr), f32(beta1), f32(beta2)
    c1, c2 = f32(1.0 - beta1), f32(1.0 - beta2)
    if kind == "sgd":
        return _csub(para, _cscale(lr32, grad)), mom1, mom2
    g2 = _abs2(grad).astype(f32)
    if kind == "rmsprop":
        sq = (b1 + c1 * g2).astype(f32)
        den = (np.sqrt(sq) + f32(1e-5)).astype(f32)
        return _csub(para, _cscale(lr32, _cdiv(grad, den))), mom1, mom2
    m1 = np.zeros_like(para) if mom1 is None else np.asarray(mom1, np.complex64)
    m2 = np.zeros(para.shape, f32) if mom2 is None else np.asarray(mom2, f32)
    m1 = (_cscale(b1, m1) + _cscale(c1, grad)).astype(np.complex64)
    if kind == "adam":
        m2 = (b2 * m2 + c2 * g2).astype(f32)
    elif kind == "yogi":
        m2 = (m2 - (c2 * _sign(m2 - g2)) * g2).astype(f32)
    else:
        raise ValueError(kind)
    den = (np.sqrt(m2) + f32(1e-5)).astype(f32)
    if kind == "adam":
        bias = f32(np.sqrt(1.0 - beta2 ** it) / (1.0 - beta1 ** it))
        upd = _cdiv(_cscale(bias, m1), den)
    else:
        upd = _cdiv(m1, den)
    return _csub(para, _cscale(lr32, upd)), m1, m2


def pupil_constraint(P, A, ac):
    """max(min(|P|, A + ac), A - ac) . sign(P) . A, sign(P) = P / |P| (0 at 0)."""
    P = np.asarray(P, np.complex64)
    A = np.asarray(A, f32)
    m = np.sqrt(P.real * P.real + P.imag * P.imag).astype(f32)
    c = np.maximum(np.minimum(m, (A + f32(ac)).astype(f32)), (A - f32(ac)).astype(f32)).astype(f32)
    with np.errstate(invalid="ignore", divide="ignore"):
        sx = np.where(m > 0, P.real / m, f32(0)).astype(f32)
        sy = np.where(m > 0, P.imag / m, f32(0)).astype(f32)
    return (((c * sx) * A) + 1j * ((c * sy) * A)).astype(np.complex64)


def solve(O0, P0, Y, led, A, pratio, epochs, batch, kind_O, kind_P, lr_O, lr_P, ac=0.1, beta1=0.9, beta2=0.999,
          prec="f64"):
    """The gradient-driven loop (MATLAB orientation, Fortran arrays; Y float
    amplitudes n x n x L, led int32 2 x L): per contiguous chunk of `batch`
    LEDs one oracle gradient, one step on O and on P, the pupil constraint.
    Adam counts iterations per optimizer from 1.  Returns (O, P)."""
    import oracle

    O = np.asfortranarray(np.asarray(O0, np.complex64))
    P = np.asfortranarray(np.asarray(P0, np.complex64))
    L = Y.shape[2]
    st = {"O": [None, None, 0], "P": [None, None, 0]}
    for _ in range(int(epochs)):
        first = 0
        while first < L:
            b = min(int(batch), L - first)
            GO, GP, _, _ = oracle.gradient(O, P, np.asfortranarray(Y[:, :, first:first + b]),
                                           np.asfortranarray(led[:, first:first + b]), pratio, prec=prec)
            for name, kind, lr in (("O", kind_O, lr_O), ("P", kind_P, lr_P)):
                s = st[name]
                s[2] += 1
                w = O if name == "O" else P
                g = GO if name == "O" else GP
                w, s[0], s[1] = step(kind, w, np.asarray(g, np.complex64), s[0], s[1], lr, beta1, beta2, s[2])
                if name == "O":
                    O = np.asfortranarray(w)
                else:
                    P = np.asfortranarray(w)
            P = np.asfortranarray(pupil_constraint(P, A, ac))
            first += b
    return O, P
