"""Pins of the warm-start network's numpy oracle (CPU only): adjoint
identities of every linear layer pair, a brute-force convolution, the
depth-to-space layout, BF16 rounding, and central finite differences of the
FP64 loss against the oracle's analytic gradient."""
import numpy as np
import pytest

import cnn_oracle as co


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


@pytest.mark.parametrize("levels", [1, 2, 3])
def test_gradient_matches_finite_differences(levels):
    """Central differences of the FP64 loss (no BF16 rounding) against the
    analytic gradient, on parameters of every conv; inputs with margins so
    no ReLU flips inside the difference step."""
    net = co.Net(levels, 32, 3, 4, bf16=False)
    n = 4 << (levels - 1)
    rng = np.random.default_rng(3 + levels)
    p = np.zeros(net.nparam)
    k = 0
    for ci, cout, _ in net.convs:
        p[k:k + cout * 9 * ci] = rng.standard_normal(cout * 9 * ci) * np.sqrt(2.0 / (9 * ci))
        k += cout * 9 * ci
        p[k:k + cout] = 0.1 * rng.standard_normal(cout)
        k += cout
    feats = rng.standard_normal((2, 3, n, n))
    targets = rng.standard_normal((2, 2, 4 * n, 4 * n))
    loss, g, _ = net.loss_and_grad(p, feats, targets)
    # probe a few weights and biases of every conv
    idx, k = [], 0
    for ci, cout, _ in net.convs:
        idx += list(rng.integers(k, k + cout * 9 * ci, 3))
        k += cout * 9 * ci
        idx.append(int(rng.integers(k, k + cout)))
        k += cout
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (net.loss_and_grad(pp, feats, targets)[0] - net.loss_and_grad(pm, feats, targets)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g).max()) + 1e-4 * abs(fd), (i, fd, g[i])
