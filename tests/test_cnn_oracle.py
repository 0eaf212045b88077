"""Pins of the warm-start network's numpy oracle (CPU only): adjoint
identities of every linear layer pair, a brute-force convolution, the
depth-to-space layout, BF16 rounding, and central finite differences of the
FP64 loss against the oracle's analytic gradient."""
import numpy as np
import pytest

import cnn_oracle as co
from helpers import texture


def test_round_bf16_ties_to_even_and_exact_values():
    # exact BF16 values stay; halfway cases round to the even mantissa
    assert co.round_bf16(np.float32(1.0)) == 1.0
    assert co.round_bf16(np.float32(1.0 + 2 ** -8)) == 1.0            # tie -> even (1.0)
    assert co.round_bf16(np.float32(1.0 + 3 * 2 ** -8)) == 1.0 + 2 ** -6  # tie -> even (up)
    assert co.round_bf16(np.float32(1.0 + 2 ** -8 + 2 ** -20)) == 1.0 + 2 ** -7
    x = np.random.default_rng(0).standard_normal(1000)
    r = co.round_bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2 ** -8)


def test_conv3x3_matches_brute_force():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 5, 6, 3))
    w = rng.standard_normal((4, 9, 3))
    b = rng.standard_normal(4)
    got = co.conv3x3(x, w, b)
    ref = np.zeros((2, 5, 6, 4))
    for bb in range(2):
        for y in range(5):
            for xx in range(6):
                for o in range(4):
                    s = b[o]
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            yy, xq = y + dy, xx + dx
                            if 0 <= yy < 5 and 0 <= xq < 6:
                                s += x[bb, yy, xq] @ w[o, 3 * (dy + 1) + dx + 1]
                    ref[bb, y, xx, o] = s
    assert np.allclose(got, ref)


@pytest.mark.parametrize("pair", ["conv_x", "conv_w", "pool", "up", "shuffle"])
def test_adjoint_identities(pair):
    """<A x, y> = <x, A^T y> for each forward map and the adjoint the backward uses."""
    rng = np.random.default_rng(2)
    B, H, W, C, K = 2, 6, 4, 3, 5
    if pair == "conv_x":
        w = rng.standard_normal((K, 9, C))
        x, y = rng.standard_normal((B, H, W, C)), rng.standard_normal((B, H, W, K))
        lhs, rhs = np.sum(co.conv3x3(x, w, None) * y), np.sum(x * co.conv3x3_input_grad(y, w))
    elif pair == "conv_w":
        x = rng.standard_normal((B, H, W, C))
        w, y = rng.standard_normal((K, 9, C)), rng.standard_normal((B, H, W, K))
        lhs, rhs = np.sum(co.conv3x3(x, w, None) * y), np.sum(w * co.conv3x3_weight_grad(y, x))
    elif pair == "pool":
        x, y = rng.standard_normal((B, H, W, C)), rng.standard_normal((B, H // 2, W // 2, C))
        lhs, rhs = np.sum(co.avgpool2(x) * y), np.sum(x * co.avgpool2_adj(y))
    elif pair == "up":
        x, y = rng.standard_normal((B, H, W, C)), rng.standard_normal((B, 2 * H, 2 * W, C))
        lhs, rhs = np.sum(co.upsample2(x) * y), np.sum(x * co.upsample2_adj(y))
    else:
        u = 4
        x, y = rng.standard_normal((B, H, W, 2 * u * u)), rng.standard_normal((B, 2, u * H, u * W))
        lhs, rhs = np.sum(co.depth_to_space(x, u) * y), np.sum(x * co.space_to_depth(y, u))
    assert abs(lhs - rhs) <= 1e-10 * max(1.0, abs(lhs))


def test_depth_to_space_layout():
    """channel c u^2 + sy u + sx of pixel (y, x) lands at plane c, (u y + sy, u x + sx)."""
    u, H, W = 4, 2, 3
    yf = np.arange(1 * H * W * 2 * u * u, dtype=np.float64).reshape(1, H, W, 2 * u * u)
    o = co.depth_to_space(yf, u)
    for c in range(2):
        for y in range(H):
            for x in range(W):
                for sy in range(u):
                    for sx in range(u):
                        assert o[0, c, u * y + sy, u * x + sx] == yf[0, y, x, c * u * u + sy * u + sx]


def test_pupil_head_adjoint_and_brute_force():
    """The pupil head's forward against explicit loops, and its backward's
    feature gradient as the adjoint of the linearised map (mean pool ->
    masked hidden layer -> output layer)."""
    rng = np.random.default_rng(9)
    B, h, w, C, H, K = 2, 3, 4, 5, co.PUPIL_HIDDEN, 3
    e = rng.standard_normal((B, h, w, C))
    W1, b1 = rng.standard_normal((H, C)), rng.standard_normal(H)
    W2, b2 = rng.standard_normal((K, H)), rng.standard_normal(K)
    pooled, hid, coef = co.pupil_head_forward(e, W1, b1, W2, b2)
    for b in range(B):
        pl = [sum(e[b, y, x, c] for y in range(h) for x in range(w)) / (h * w) for c in range(C)]
        hd = [max(0.0, b1[j] + sum(W1[j, c] * pl[c] for c in range(C))) for j in range(H)]
        for k in range(K):
            assert abs(coef[b, k] - (b2[k] + sum(W2[k, j] * hd[j] for j in range(H)))) <= 1e-9 * (1 + abs(coef[b, k]))
    # <J de, g> = <de, J^T g> with J the derivative of coef in e (ReLU mask fixed)
    g = rng.standard_normal((B, K))
    de = rng.standard_normal(e.shape)
    mask = hid > 0
    J_de = ((de.mean(axis=(1, 2)) @ W1.T) * mask) @ W2.T
    *_, gpool = co.pupil_head_backward(g, pooled, hid, W1, W2, h * w)
    JT_g = np.broadcast_to(gpool[:, None, None, :], e.shape)
    assert abs(np.sum(J_de * g) - np.sum(de * JT_g)) <= 1e-10 * max(1.0, abs(np.sum(J_de * g)))


@pytest.mark.parametrize("levels,pk", [(1, 0), (2, 0), (3, 0), (1, 3), (2, 5)])
def test_gradient_matches_finite_differences(levels, pk):
    """Central differences of the FP64 loss (no BF16 rounding) against the
    analytic gradient, on parameters of every conv and of the pupil head;
    inputs with margins so no ReLU flips inside the difference step."""
    net = co.Net(levels, 32, 3, 4, bf16=False, pupil_terms=pk, pupil_weight=0.7)
    n = 4 << (levels - 1)
    rng = np.random.default_rng(3 + levels + 10 * pk)
    p = np.zeros(net.nparam)
    k = 0
    for ci, cout, _ in net.convs:
        p[k:k + cout * 9 * ci] = rng.standard_normal(cout * 9 * ci) * np.sqrt(2.0 / (9 * ci))
        k += cout * 9 * ci
        p[k:k + cout] = 0.1 * rng.standard_normal(cout)
        k += cout
    p[k:] = 0.2 * rng.standard_normal(net.nparam - k)
    feats = rng.standard_normal((2, 3, n, n))
    targets = rng.standard_normal((2, 2, 4 * n, 4 * n))
    zt = rng.normal(0.0, 0.4, (2, pk)) if pk else None
    loss, g, _ = net.loss_and_grad(p, feats, targets, zt)
    # probe a few weights and biases of every conv and of the pupil head's layers
    idx, k = [], 0
    sizes = []
    for ci, cout, _ in net.convs:
        sizes += [cout * 9 * ci, cout]
    if pk:
        sizes += [co.PUPIL_HIDDEN * net.cb, co.PUPIL_HIDDEN, pk * co.PUPIL_HIDDEN, pk]
    for size in sizes:
        idx += list(rng.integers(k, k + size, 3 if size > 64 else 1))
        k += size
    assert k == net.nparam
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (net.loss_and_grad(pp, feats, targets, zt)[0] - net.loss_and_grad(pm, feats, targets, zt)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g).max()) + 1e-4 * abs(fd), (i, fd, g[i])


def test_parallax_shift_zero_and_known_shifts():
    """Identical planes give exactly zero; a texture translated by a known
    sub-pixel amount (Fourier shift theorem, so the truth does not depend on
    any interpolation) is recovered within 0.02 px; swapping the planes
    negates the estimate; an affine change of brightness does not matter."""
    ref = texture(96, 1)
    assert np.array_equal(co.parallax_shift(ref, ref, 4), [0.0, 0.0])
    for d in ((0.3, -0.45), (-0.8, 0.1), (1.6, 0.7)):
        img = texture(96, 1, d)
        est = co.parallax_shift(ref, img, 6)
        assert np.all(np.abs(est - d) <= 0.02), (d, est)
        back = co.parallax_shift(img, ref, 6)
        assert np.all(np.abs(back + est) <= 0.02), (est, back)
        assert np.allclose(co.parallax_shift(ref, 3.0 * img + 7.0, 6), est, atol=1e-9)


def test_parallax_bilinear_matches_brute_force():
    """The oracle's clamped bilinear sampler against explicit loops."""
    rng = np.random.default_rng(4)
    q = rng.standard_normal((5, 7))
    ys = rng.uniform(-1.5, 5.5, 40)
    xs = rng.uniform(-1.5, 7.5, 40)
    got = co._bilinear_clamped(q, ys, xs)
    for k in range(40):
        y, x = min(max(ys[k], 0.0), 4.0), min(max(xs[k], 0.0), 6.0)
        y0, x0 = min(int(np.floor(y)), 3), min(int(np.floor(x)), 5)
        fy, fx = y - y0, x - x0
        ref = ((1 - fy) * (1 - fx) * q[y0, x0] + (1 - fy) * fx * q[y0, x0 + 1] + fy * (1 - fx) * q[y0 + 1, x0]
               + fy * fx * q[y0 + 1, x0 + 1])
        assert abs(got[k] - ref) <= 1e-12
