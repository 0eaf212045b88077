"""Warm-start network on the GPU (fpm_cnn_*): forward and backward kernels
against the numpy oracle (BF16 rounding emulated at the library's storage
points), training, and what the trained network gains the solve."""
import numpy as np
import pytest

import cnn_oracle
import fpm_amd
from fpm_amd import warmstart
from helpers import rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _net(levels=3, n=16, batch=2, seed=1, random_head=True):
    """Network as created (He-normal convs, zero head) with, by default, a
    random head, so that every layer's output and gradient is non-trivial."""
    net = warmstart.WarmStartNet(n, 4, levels=levels, base=32, in_channels=26, max_batch=batch, seed=seed)
    if random_head:
        p = net.get_params()
        ci, co, _ = cnn_oracle.layer_shapes(levels, 32, 26, 4)[-1]
        k = p.size - (co * 9 * ci + co)
        p[k:k + co * 9 * ci] = np.random.default_rng(seed).standard_normal(co * 9 * ci) * np.sqrt(2.0 / (9 * ci))
        net.set_params(p)
    return net


def _inputs(n, batch, seed):
    rng = np.random.default_rng(seed)
    f = rng.random((batch, 26, n, n)).astype(np.float32)
    t = rng.standard_normal((batch, 2, 4 * n, 4 * n)).astype(np.float32)
    return f, t


@pytest.mark.parametrize("levels,n", [(1, 8), (3, 16), (4, 32)])
def test_cnn_forward_matches_oracle(levels, n):
    """Network output vs the oracle with the same parameters: relative L2
    within 1e-2 (BF16 activations: an FP32 accumulation that lands on the
    other side of a BF16 rounding boundary moves a value by 2^-8)."""
    net = _net(levels, n)
    p = net.get_params()
    f, _ = _inputs(n, 2, levels)
    got = net.forward(torch.from_numpy(f).cuda()).cpu().numpy()
    ref, _ = cnn_oracle.Net(levels, 32, 26, 4, bf16=True).forward(p.astype(np.float64), f.astype(np.float64))
    print(f"[cnn-fwd] levels {levels} n {n}: rel-L2 {rel_l2(got, ref):.2e}", flush=True)
    assert rel_l2(got, ref) <= 1e-2
    net.close()


@pytest.mark.parametrize("levels,n", [(1, 8), (3, 16), (4, 32)])
def test_cnn_gradient_matches_oracle(levels, n):
    """Loss within 1e-3 and the gradient of every conv's weights and biases
    within 5e-2 relative L2 of the oracle's (BF16 gradients entering each
    product, FP32 accumulation against FP64)."""
    net = _net(levels, n)
    p = net.get_params()
    f, t = _inputs(n, 2, 10 + levels)
    loss, g = net.gradient(torch.from_numpy(f).cuda(), torch.from_numpy(t).cuda())
    onet = cnn_oracle.Net(levels, 32, 26, 4, bf16=True)
    lref, gref, _ = onet.loss_and_grad(p.astype(np.float64), f.astype(np.float64), t.astype(np.float64))
    assert abs(loss - lref) <= 1e-3 * lref
    k = 0
    worst = 0.0
    for ci, cout, _ in onet.convs:
        for size in (cout * 9 * ci, cout):
            e = rel_l2(g[k:k + size], gref[k:k + size])
            worst = max(worst, e)
            k += size
    print(f"[cnn-grad] levels {levels} n {n}: loss {loss:.6f} vs {lref:.6f}, worst per-tensor rel-L2 {worst:.2e}, "
          f"whole {rel_l2(g, gref):.2e}", flush=True)
    assert rel_l2(g, gref) <= 2e-2
    assert worst <= 5e-2
    net.close()


def test_cnn_training_fits_a_batch():
    """Adam steps through the library's own backward kernels drive the loss
    of a fixed batch down by 5x in 150 steps (the gradient is a descent
    direction end to end).  The target is learnable by the network: the
    first two feature planes upsampled 4x (nearest), centred."""
    net = _net(3, 16, seed=3, random_head=False)
    f, _ = _inputs(16, 2, 21)
    t = np.repeat(np.repeat(f[:, :2] - 0.5, 4, axis=2), 4, axis=3).astype(np.float32)
    F, T = torch.from_numpy(f).cuda(), torch.from_numpy(np.ascontiguousarray(t)).cuda()
    l0 = net.train_step(F, T, 2e-3)
    for _ in range(150):
        l = net.train_step(F, T, 2e-3)
    print(f"[cnn-train] loss {l0:.4f} -> {l:.4f}", flush=True)
    assert l < 0.2 * l0
    net.close()


def test_cnn_save_load_roundtrip(tmp_path):
    """Parameters survive save/load (plain npz arrays) and give the same output."""
    net = _net(2, 16, seed=4)
    f, _ = _inputs(16, 2, 30)
    F = torch.from_numpy(f).cuda()
    out = net.forward(F).cpu().numpy()
    path = str(tmp_path / "net.npz")
    net.save(path)
    net2 = warmstart.WarmStartNet.load(path)
    assert np.array_equal(net2.get_params(), net.get_params())
    assert np.array_equal(net2.forward(F).cpu().numpy(), out)
    net.close()
    net2.close()


def _phase_aligned_rel_l2(a, b):
    """min over a global phase factor of ||a e^{i phi} - b|| / ||b||"""
    a = a.astype(np.complex128).ravel()
    b = b.astype(np.complex128).ravel()
    c = np.vdot(a, b)
    ph = c / abs(c) if abs(c) > 0 else 1.0
    return float(np.linalg.norm(a * ph - b) / np.linalg.norm(b))


def test_warm_start_gains_over_plain_init():
    """Train a small network (3 levels, 32 x 32 features -> 128 x 128 object,
    7 x 7 LEDs) on 384 synthetic objects, then on 4 held-out objects compare
    the solve from the network's object with the solve from the reference's
    init (upsampled centre image): the initial error against the true object
    and the loss after 4 epochs must both be lower with the network."""
    n, r, grid = 32, 4, 7
    net = warmstart.WarmStartNet(n, r, levels=3, base=32, max_batch=8, seed=5)
    F, T = warmstart.make_training_set(384, n, r, seed0=5000, led_grid=grid)
    hist = warmstart.train(net, F, T, epochs=12, batch=8, lr=2e-3, seed=1)
    print(f"[warmstart] training loss {hist[0]:.4f} -> {hist[-1]:.4f}", flush=True)
    assert hist[-1] < hist[0]
    e_plain, e_warm, l_plain, l_warm = [], [], [], []
    for k in range(4):
        feats, target, ds = warmstart.make_sample(n, r, 90000 + k, led_grid=grid)
        O_true = ds["O_true_dev"].cpu().numpy()
        O_net = warmstart.warm_start_object(net, feats, ds["O0_dev"].real)
        for O0, errs, losses in ((ds["O0_dev"], e_plain, l_plain), (O_net, e_warm, l_warm)):
            errs.append(_phase_aligned_rel_l2(O0.cpu().numpy(), O_true))
            Od, Pd = torch.empty_like(O0), torch.empty_like(ds["P0_dev"])
            torch.cuda.synchronize()
            lossv = fpm_amd.solve_device(O0.data_ptr(), ds["P0_dev"].data_ptr(), ds["raw_dev"].data_ptr(), True,
                                         ds["led_dev"].data_ptr(), ds["A_dev"].data_ptr(), (n * r, n * r), (n, n),
                                         ds["L"], r, 4, 26, (2e-3, 0.999, 1e-4), Od.data_ptr(), Pd.data_ptr(),
                                         loss=True)
            losses.append(lossv[-1])
    print(f"[warmstart] init error plain {np.mean(e_plain):.3f} net {np.mean(e_warm):.3f}; "
          f"loss after 4 epochs plain {np.mean(l_plain):.4e} net {np.mean(l_warm):.4e}", flush=True)
    assert np.mean(e_warm) < np.mean(e_plain)
    assert np.mean(l_warm) < np.mean(l_plain)
    net.close()
