"""Warm-start network on the GPU (fpm_cnn_*): forward and backward kernels
against the numpy oracle (BF16 rounding emulated at the library's storage
points), training, and what the trained network gains the solve."""
import numpy as np
import pytest

import cnn_oracle
import fpm_amd
from fpm_amd import warmstart
from helpers import rel_l2, texture

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _net(levels=3, n=16, batch=2, seed=1, random_head=True, pupil_terms=0):
    """Network as created (He-normal convs, zero head, zero pupil output
    layer) with, by default, random heads, so that every layer's output and
    gradient is non-trivial."""
    net = warmstart.WarmStartNet(n, 4, levels=levels, base=32, in_channels=26, max_batch=batch, seed=seed,
                                 pupil_terms=pupil_terms, pupil_weight=0.5)
    if random_head:
        p = net.get_params()
        rng = np.random.default_rng(seed)
        onet = cnn_oracle.Net(levels, 32, 26, 4, pupil_terms=pupil_terms)
        ci, co, _ = onet.convs[-1]
        k = onet.nconv - (co * 9 * ci + co)
        p[k:k + co * 9 * ci] = rng.standard_normal(co * 9 * ci) * np.sqrt(2.0 / (9 * ci))
        if pupil_terms:
            h = cnn_oracle.PUPIL_HIDDEN
            k2 = onet.nconv + h * onet.cb + h
            p[k2:k2 + pupil_terms * h] = rng.standard_normal(pupil_terms * h) * np.sqrt(1.0 / h)
            p[onet.nconv + h * onet.cb:k2] = 0.1 * rng.standard_normal(h)  # b1: some hidden units on
        net.set_params(p)
    return net


def _tensor_sizes(onet):
    """sizes of the parameter tensors in vector order"""
    sizes = []
    for ci, cout, _ in onet.convs:
        sizes += [cout * 9 * ci, cout]
    if onet.K:
        h = cnn_oracle.PUPIL_HIDDEN
        sizes += [h * onet.cb, h, onet.K * h, onet.K]
    return sizes


def _inputs(n, batch, seed):
    rng = np.random.default_rng(seed)
    f = rng.random((batch, 26, n, n)).astype(np.float32)
    t = rng.standard_normal((batch, 2, 4 * n, 4 * n)).astype(np.float32)
    return f, t


# (levels, n, pupil terms); n = 64 runs the 64-pixel-row patch convolution on its first level
CASES = [(1, 8, 0), (3, 16, 0), (4, 32, 0), (2, 64, 0), (1, 8, 5), (3, 16, 5), (4, 32, 5)]


@pytest.mark.parametrize("levels,n,pk", CASES)
def test_cnn_forward_matches_oracle(levels, n, pk):
    """Network output vs the oracle with the same parameters: relative L2
    within 1e-2 (BF16 activations: an FP32 accumulation that lands on the
    other side of a BF16 rounding boundary moves a value by 2^-8); the pupil
    coefficients (FP32 head on BF16 features) within 1e-2 too."""
    net = _net(levels, n, pupil_terms=pk)
    p = net.get_params()
    f, _ = _inputs(n, 2, levels)
    res = net.forward(torch.from_numpy(f).cuda(), pupil=pk > 0)
    got = (res[0] if pk else res).cpu().numpy()
    ref, cache = cnn_oracle.Net(levels, 32, 26, 4, bf16=True, pupil_terms=pk).forward(p.astype(np.float64),
                                                                                        f.astype(np.float64))
    msg = f"[cnn-fwd] levels {levels} n {n} pupil {pk}: rel-L2 {rel_l2(got, ref):.2e}"
    if pk:
        ec = rel_l2(res[1].cpu().numpy(), cache["coef"])
        msg += f", coefficients {ec:.2e}"
        assert ec <= 1e-2
    print(msg, flush=True)
    assert rel_l2(got, ref) <= 1e-2
    net.close()


@pytest.mark.parametrize("levels,n,pk", CASES)
def test_cnn_gradient_matches_oracle(levels, n, pk):
    """Loss within 1e-3 and the gradient of every parameter tensor (conv
    weights and biases, pupil head layers) within 5e-2 relative L2 of the
    oracle's (BF16 gradients entering each conv product, FP32 accumulation
    against FP64)."""
    net = _net(levels, n, pupil_terms=pk)
    p = net.get_params()
    f, t = _inputs(n, 2, 10 + levels)
    z = np.random.default_rng(20 + levels).normal(0.0, 0.4, (2, pk)).astype(np.float32) if pk else None
    loss, g = net.gradient(torch.from_numpy(f).cuda(), torch.from_numpy(t).cuda(),
                           torch.from_numpy(z).cuda() if pk else None)
    onet = cnn_oracle.Net(levels, 32, 26, 4, bf16=True, pupil_terms=pk, pupil_weight=0.5)
    lref, gref, _ = onet.loss_and_grad(p.astype(np.float64), f.astype(np.float64), t.astype(np.float64),
                                       z.astype(np.float64) if pk else None)
    assert abs(loss - lref) <= 1e-3 * lref
    k = 0
    worst = 0.0
    for size in _tensor_sizes(onet):
        e = rel_l2(g[k:k + size], gref[k:k + size])
        worst = max(worst, e)
        k += size
    assert k == g.size
    print(f"[cnn-grad] levels {levels} n {n} pupil {pk}: loss {loss:.6f} vs {lref:.6f}, worst per-tensor rel-L2 "
          f"{worst:.2e}, whole {rel_l2(g, gref):.2e}", flush=True)
    assert rel_l2(g, gref) <= 2e-2
    assert worst <= 5e-2
    net.close()


def test_cnn_training_fits_a_batch():
    """Adam steps through the library's own backward kernels drive the loss
    of a fixed batch down by 5x in 150 steps (the gradient is a descent
    direction end to end), and the pupil head's coefficient error down by
    5x.  The targets are learnable by the network: the first two feature
    planes upsampled 4x (nearest), centred; coefficients = 4 x (the mean of
    feature planes 2..6 - 0.5)."""
    net = _net(3, 16, seed=3, random_head=False, pupil_terms=5)
    f, _ = _inputs(16, 2, 21)
    f[1, 2:7] *= 0.5  # different plane means in the two samples
    t = np.repeat(np.repeat(f[:, :2] - 0.5, 4, axis=2), 4, axis=3).astype(np.float32)
    z = (4.0 * (f[:, 2:7].mean(axis=(2, 3)) - 0.5)).astype(np.float32)
    F, T = torch.from_numpy(f).cuda(), torch.from_numpy(np.ascontiguousarray(t)).cuda()
    Z = torch.from_numpy(z).cuda()
    c0 = np.linalg.norm(net.forward(F, pupil=True)[1].cpu().numpy() - z)
    l0 = net.train_step(F, T, 2e-3, Z)
    for _ in range(150):
        l = net.train_step(F, T, 2e-3, Z)
    c = np.linalg.norm(net.forward(F, pupil=True)[1].cpu().numpy() - z)
    print(f"[cnn-train] loss {l0:.4f} -> {l:.4f}, coefficient error {c0:.4f} -> {c:.4f}", flush=True)
    assert l < 0.2 * l0
    assert c < 0.2 * c0
    net.close()


def test_cnn_save_load_roundtrip(tmp_path):
    """Parameters and the pupil head survive save/load (plain npz arrays) and
    give the same output."""
    net = _net(2, 16, seed=4, pupil_terms=5)
    f, _ = _inputs(16, 2, 30)
    F = torch.from_numpy(f).cuda()
    out = net.forward(F).cpu().numpy()
    path = str(tmp_path / "net.npz")
    net.save(path)
    net2 = warmstart.WarmStartNet.load(path)
    assert net2.desc.pupil_terms == 5 and net2.desc.pupil_weight == net.desc.pupil_weight
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
    7 x 7 LEDs, the generator's fixed pupil) on 384 synthetic objects, then
    on 4 held-out objects compare the solve from the network's object with
    the solve from the reference's init (upsampled centre image): the
    initial error against the true object and the loss after 4 epochs must
    both be lower with the network."""
    n, r, grid = 32, 4, 7
    net = warmstart.WarmStartNet(n, r, levels=3, base=32, max_batch=8, seed=5)
    F, T, _ = warmstart.make_training_set(384, n, r, seed0=5000, led_grid=grid, pupil_sigma=None)
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


def test_parallax_pupil_beats_the_ideal_pupil():
    """Fit the parallax pupil map on 256 synthetic objects (64 x 64 images,
    upsample 2, 19 x 19 LEDs, pupils with 5 Zernike terms ~ N(0, 0.4) rad);
    on 8 held-out objects the predicted coefficients' RMS error is below
    0.6x that of the ideal pupil (all zero), and a 3-epoch solve from the
    predicted pupil ends at a lower loss than from the ideal pupil."""
    n, r = 64, 2
    F, _, Z = warmstart.make_training_set(256, n, r, seed0=7000)
    _, _, nai = fpm_amd.led_offsets(n, r)
    pupil = warmstart.ParallaxPupil(warmstart.brightfield_planes(nai))
    fit = pupil.fit(F, Z)
    del F, Z
    err, ideal, l_est, l_ideal = [], [], [], []
    for k in range(8):
        feats, _, ds = warmstart.make_sample(n, r, 95000 + k, pupil_sigma=warmstart.PUPIL_SIGMA)
        coef = pupil.predict(feats[None])
        err.append(float(np.sum((coef[0].cpu().numpy() - ds["pupil_coeffs"]) ** 2)))
        ideal.append(float(np.sum(ds["pupil_coeffs"] ** 2)))
        if k < 4:
            P_est = warmstart.pupil_from_coeffs(coef[0], ds["A_dev"])
            for P0, losses in ((P_est, l_est), (ds["P0_dev"], l_ideal)):
                _, lv, _, _ = warmstart._timed_solve(ds, ds["O0_dev"], 3, 26, (5e-4, 0.999, 1e-4), fpm_amd.FFT_F32,
                                                     0, P0)
                losses.append(lv[-1])
    e, z = np.sqrt(np.mean(err)), np.sqrt(np.mean(ideal))
    print(f"[parallax-pupil] fit RMS {fit:.3f}; held-out coefficient error {e:.3f} vs ideal pupil {z:.3f}; "
          f"loss after 3 epochs {np.mean(l_est):.4e} vs {np.mean(l_ideal):.4e}", flush=True)
    assert e < 0.6 * z
    assert np.mean(l_est) < np.mean(l_ideal)


def test_parallax_shifts_match_oracle():
    """fpm_parallax_shifts against the oracle on textures translated by known
    sub-pixel amounts (2 samples of 4 planes, plane 0 the reference; one
    plane with an affine brightness change): within 2e-3 px of the oracle
    and 0.02 px of the true shifts."""
    n = 80
    shifts = [[(0.3, -0.45), (-0.8, 0.1), (1.4, 0.6)], [(0.0, 0.0), (-0.2, -1.1), (0.55, 0.25)]]
    planes = np.zeros((2, 4, n, n), np.float32)
    for b in range(2):
        planes[b, 0] = texture(n, 10 + b)
        for j, d in enumerate(shifts[b]):
            planes[b, 1 + j] = texture(n, 10 + b, d)
    planes[1, 3] = 2.5 * planes[1, 3] + 3.0
    got = warmstart.parallax_shifts(torch.from_numpy(planes).cuda(), [0, 1, 2, 3], iters=6).cpu().numpy()
    worst_o, worst_t = 0.0, 0.0
    for b in range(2):
        for j, d in enumerate(shifts[b]):
            ref = cnn_oracle.parallax_shift(planes[b, 0], planes[b, 1 + j], 6)
            g = got[b, 2 * j:2 * j + 2]
            worst_o = max(worst_o, float(np.abs(g - ref).max()))
            worst_t = max(worst_t, float(np.abs(g - np.array(d)).max()))
    print(f"[parallax] worst |gpu - oracle| {worst_o:.2e} px, |gpu - truth| {worst_t:.2e} px", flush=True)
    assert worst_o <= 2e-3
    assert worst_t <= 0.02
