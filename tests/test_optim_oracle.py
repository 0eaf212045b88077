"""The optimizer oracle (oracle/optim_oracle.py) pinned by what the
mathematics fixes, independently of its own formulas: Adam's bias-corrected
first step has magnitude lr, Yogi's first step from zero moments is
(1 - b1)/sqrt(1 - b2) lr and a tie leaves its second moment alone, the
RMSprop as written keeps no state, SGD is the plain step, and the pupil
constraint clamps the modulus into the band and keeps the phase."""
import numpy as np

import optim_oracle as oo

RNG = np.random.default_rng(3)


def _grad(shape, scale=1.0):
    return (scale * (RNG.standard_normal(shape) + 1j * RNG.standard_normal(shape))).astype(np.complex64)


def test_adam_first_step_has_magnitude_lr():
    p0, g = _grad((64, 64)), _grad((64, 64))
    lr = 1e-3
    p1, m1, m2 = oo.step("adam", p0, g, lr=lr, beta1=0.9, beta2=0.999, it=1)
    d = np.abs(p1.astype(np.complex128) - p0)
    big = np.abs(g) > 0.5  # |g| >> 1e-5 / sqrt(1 - b2) = 3.2e-4
    assert np.allclose(d[big], lr, rtol=2e-3)
    # direction of the step = direction of the gradient
    cosang = np.real((p0 - p1.astype(np.complex128)) * np.conj(g)) / (d * np.abs(g))
    assert np.all(cosang[big] > 1 - 1e-4)
    # the moments of step 1 are (1 - b) times the gradient terms
    assert np.allclose(m1, 0.1 * g, rtol=1e-6)
    assert np.allclose(m2, 0.001 * np.abs(g) ** 2, rtol=1e-5)


def test_adam_bias_correction_uses_the_iteration():
    """At iteration t with constant g the moments are (1 - b^t) g and
    (1 - b2^t)|g|^2, so the corrected step stays lr."""
    p, g = _grad((32, 32)), _grad((32, 32))
    m1 = m2 = None
    lr = 1e-3
    for t in range(1, 6):
        q, m1, m2 = oo.step("adam", p, g, m1, m2, lr=lr, it=t)
        d = np.abs(q.astype(np.complex128) - p)
        big = np.abs(g) > 0.5
        assert np.allclose(d[big], lr, rtol=3e-3), t
        p = q


def test_yogi_first_step_and_tie():
    p0, g = _grad((64, 64)), _grad((64, 64))
    lr = 1e-3
    p1, m1, m2 = oo.step("yogi", p0, g, lr=lr)
    d = np.abs(p1.astype(np.complex128) - p0)
    big = np.abs(g) > 0.5
    assert np.allclose(d[big], lr * 0.1 / np.sqrt(0.001), rtol=3e-3)
    # |g|^2 above m2 raises m2 by (1 - b2)|g|^2, below lowers it by the same
    g_small = (0.5 * g).astype(np.complex64)
    _, _, m2b = oo.step("yogi", p1, g_small, m1, m2, lr=lr)
    assert np.all(m2b <= m2) or np.all(np.abs(g_small) ** 2 >= m2)
    # a tie (m2 == |g|^2 in single) leaves m2 unchanged: MATLAB sign(0) = 0
    g2 = (np.sqrt(g.real * g.real + g.imag * g.imag) ** 2).astype(np.float32)
    _, _, m2t = oo.step("yogi", p0, g, np.zeros_like(p0), g2, lr=lr)
    assert np.array_equal(m2t, g2)


def test_rmsprop_as_written_keeps_no_state():
    # p starts at 0 so that steps of lr |g| ~ 1e-7 are resolved in single
    p0, g = np.zeros((32, 32), np.complex64), _grad((32, 32), 1e-4)
    lr = 1e-3
    p1, _, _ = oo.step("rmsprop", p0, g, lr=lr, beta1=0.999)
    p2, _, _ = oo.step("rmsprop", p1, g, lr=lr, beta1=0.999)
    d1 = p1.astype(np.complex128) - p0
    d2 = p2.astype(np.complex128) - p1
    assert np.allclose(d1, d2, rtol=1e-3, atol=1e-9)
    # small |g|: squave -> beta1, step -> lr g / sqrt(beta1)
    assert np.allclose(-d1, lr * g / np.sqrt(0.999), rtol=1e-3, atol=1e-10)


def test_sgd_is_the_plain_step():
    p0, g = _grad((16, 16)), _grad((16, 16))
    p1, _, _ = oo.step("sgd", p0, g, lr=0.25)  # exact in binary
    assert np.array_equal(p1, (p0 - 0.25 * g).astype(np.complex64))


def test_pupil_constraint_band_and_phase():
    n = 64
    yy, xx = np.mgrid[:n, :n]
    A = ((yy - n / 2) ** 2 + (xx - n / 2) ** 2 <= (n / 4) ** 2).astype(np.float32)
    P = _grad((n, n), 0.8)
    ac = 0.1
    Q = oo.pupil_constraint(P, A, ac)
    m = np.abs(Q.astype(np.complex128))
    inside = A > 0
    assert np.all(m[inside] >= 1 - ac - 1e-6) and np.all(m[inside] <= 1 + ac + 1e-6)
    assert np.all(Q[~inside] == 0)
    ph = np.angle(Q[inside].astype(np.complex128) * np.conj(P[inside]))
    assert np.max(np.abs(ph)) < 1e-6
    # a modulus already in the band is kept
    band = inside & (np.abs(P) > 1 - ac) & (np.abs(P) < 1 + ac)
    assert np.allclose(Q[band], P[band], rtol=1e-6)
    # zero stays zero (sign(0) = 0)
    P[n // 2, n // 2] = 0
    assert oo.pupil_constraint(P, A, ac)[n // 2, n // 2] == 0


def test_loop_first_chunk_is_one_step():
    """The oracle loop over one chunk = one gradient (oracle.gradient), one
    SGD step each, the constraint: checks the loop's slicing and order."""
    import oracle
    from helpers import margin_problem

    O, P, Y, led = margin_problem(16, 2, 3)
    n = P.shape[0]
    A = np.ones((n, n), np.float32)
    L = Y.shape[2]
    Oo, Po = oo.solve(O, P, Y, led, A, 2, 1, L, "sgd", "sgd", 1e-3, 2e-3, ac=10.0)
    GO, GP, _, _ = oracle.gradient(O, P, Y, led, 2)
    O1 = (O - np.float32(1e-3) * GO.astype(np.complex64)).astype(np.complex64)
    P1 = oo.pupil_constraint((P - np.float32(2e-3) * GP.astype(np.complex64)).astype(np.complex64), A, 10.0)
    assert np.allclose(Oo, O1, rtol=1e-6, atol=1e-7)
    assert np.allclose(Po, P1, rtol=1e-6, atol=1e-7)
