"""The reference's MATLAB-side optimizers and pupil constraint on the GPU
(fpm_optimizer_step, fpm_pupil_constraint) against the single-precision
oracle (oracle/optim_oracle.py), and the gradient-driven solve
(fpm_amd.optimizers.solve_with_optimizers: fused gradient + optimizer steps
+ constraint per minibatch) against the oracle loop on the FP64 gradient.
Tolerances: the steps use the same float operations in the same order with
contraction off, so they agree to an ulp (rtol 1e-6); the loop compares a
GPU FP32 gradient with the FP64 oracle over 2 epochs of normalised steps on
inputs whose loss-stencil signs all have a margin, rel-L2 <= 1e-6 (measured
1.5e-8 / 6.5e-8)."""
import numpy as np
import pytest

import optim_oracle as oo
from fpm_amd import optimizers as fo
from fpm_amd import synth
from helpers import margin_problem, rel_l2

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _data(n, seed):
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64)
    gs = [(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64) for _ in range(3)]
    for g in gs:
        g[0, :8] = 0  # zero gradients
    return p, gs


def _close(a, b):
    scale = float(np.max(np.abs(b)))
    np.testing.assert_allclose(a, b, rtol=1e-6, atol=1e-6 * scale)


@pytest.mark.parametrize("kind", ["adam", "yogi", "rmsprop", "sgd"])
def test_optimizer_steps_match_the_oracle(kind):
    n = 96
    p, gs = _data(n, 7)
    make = {"adam": lambda: fo.optimizer_adam(0, 0, 0.9, 0.999, 1e-3),
            "yogi": lambda: fo.optimizer_yogi(0, 0, 0.9, 0.999, 2e-3),
            "rmsprop": lambda: fo.optimizer_RMSprop(0, 0, 0.999, 0, False, 2e-3),
            "sgd": lambda: fo.optimizer_sgd(1e-2)}[kind]
    lr = {"adam": 1e-3, "yogi": 2e-3, "rmsprop": 2e-3, "sgd": 1e-2}[kind]
    b1 = 0.999 if kind == "rmsprop" else 0.9
    opt = make()
    pd = torch.from_numpy(p).cuda()
    ref, m1, m2 = p, None, None
    for it, g in enumerate(gs, start=1):
        opt.step(pd, torch.from_numpy(g).cuda())
        ref, m1, m2 = oo.step(kind, ref, g, m1, m2, lr=lr, beta1=b1, beta2=0.999, it=it)
        torch.cuda.synchronize()
        _close(pd.cpu().numpy(), ref)
        if kind in ("adam", "yogi"):
            _close(opt.mom1.cpu().numpy(), m1)
            _close(opt.mom2.cpu().numpy(), m2)


def test_yogi_tie_keeps_the_second_moment():
    """m2 == |g|^2 in single: MATLAB sign(0) = 0, m2 unchanged on the GPU too."""
    n = 64
    p, gs = _data(n, 8)
    g = gs[0]
    g2 = (np.sqrt(g.real * g.real + g.imag * g.imag) ** 2).astype(np.float32)
    opt = fo.optimizer_yogi(0, 0, 0.9, 0.999, 1e-3)
    pd = torch.from_numpy(p).cuda()
    opt._ensure(pd)
    opt.mom2.copy_(torch.from_numpy(g2))
    opt.step(pd, torch.from_numpy(g).cuda())
    torch.cuda.synchronize()
    assert np.array_equal(opt.mom2.cpu().numpy(), g2)


def test_pupil_constraint_matches_the_oracle():
    n = 128
    A, _ = synth.pupil(n)
    rng = np.random.default_rng(9)
    P = (0.9 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))).astype(np.complex64)
    P[5, 5] = 0
    Pd = torch.from_numpy(np.ascontiguousarray(P)).cuda()
    fo.pupil_constraint(Pd, torch.from_numpy(np.ascontiguousarray(A.astype(np.float32))).cuda(), 0.1)
    torch.cuda.synchronize()
    _close(Pd.cpu().numpy(), oo.pupil_constraint(P, A.astype(np.float32), 0.1))


def test_gradient_driven_solve_matches_the_oracle_loop():
    """2 epochs over 25 LEDs in chunks of 10 (a ragged last chunk), Adam on
    O and Yogi on P (the reference's commented-out alternative pairs these
    with RMSprop), constraint band 0.1, FP64 oracle gradient."""
    n, r = 32, 4
    O, P, Y, led = margin_problem(n, r, 5)
    A, _ = synth.pupil(n)
    A = A.astype(np.float32)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a.T)).cuda()  # noqa: E731
    Yd = torch.from_numpy(np.ascontiguousarray(Y.transpose(2, 1, 0).astype(np.float32))).cuda()
    Od, Pd, losses = fo.solve_with_optimizers(dev(O.astype(np.complex64)), dev(P.astype(np.complex64)), Yd,
                                              dev(led.astype(np.int32)), dev(A), r, 2, 10,
                                              fo.optimizer_adam(0, 0, 0.9, 0.999, 1e-3),
                                              fo.optimizer_yogi(0, 0, 0.9, 0.999, 2e-3), 0.1, loss=True)
    Og, Pg = Od.cpu().numpy().T, Pd.cpu().numpy().T
    Oo, Po = oo.solve(O, P, Y, led, A, r, 2, 10, "adam", "yogi", 1e-3, 2e-3, ac=0.1, prec="f64")
    eo, ep = rel_l2(Og, Oo), rel_l2(Pg, Po)
    print(f"[optim] gradient-driven solve vs oracle loop: O {eo:.2e} P {ep:.2e}; losses {losses[:3]}", flush=True)
    assert len(losses) == 6 and all(np.isfinite(losses))
    assert eo <= 1e-6 and ep <= 1e-6
