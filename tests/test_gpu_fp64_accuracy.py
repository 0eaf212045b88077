"""Accuracy against the FP64 oracle at the workloads the benchmark runs:
the full 2048^2 x 361 -> 16384^2 gradient of one minibatch (and the 512^2 x
361 -> 4096^2 one), and a short solve from an aberrated-pupil start at the
warm-start size.  With real data some loss-stencil sign decisions lie within
FP32 rounding of a tie, so every FP32 computation (the GPU's and the
oracle's own FP32 build) takes a few of them differently from FP64 and its
gradient lands about 1e-3 from FP64's.  Two bars separate the causes: (1)
sign-synchronised -- the FP64 oracle given the GPU's own sign decisions
(fpm_opts.stencil_signs) -- the GPU gradient is within 1e-5 of FP64, so the
arithmetic (transforms, products, sums) meets the 1e-5 bar at full scale;
(2) unsynchronised, both GPU engines stay as close to FP64 as the FP32
oracle does (a multiple of its distance, plus 1e-5, below a cap)."""
import numpy as np
import pytest

import fpm_amd
import oracle
from fpm_amd import synth, warmstart
from helpers import margin_problem, rect_margin_problem, rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _close_enough(name, e, ref, factor, cap):
    for k in e:
        assert e[k] <= factor[k] * ref[k] + 1e-5 and e[k] <= cap, (name, k, e[k], ref[k])


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _host(t):
    """device layout (row-major buffer of a column-major array) -> MATLAB-oriented host array"""
    return np.asfortranarray(t.cpu().numpy().T)


def _prepared(ds, count):
    raw = ds["raw_dev"][:count].cpu().numpy()
    mn, mx = ds["raw_min"], ds["raw_max"]
    Y = np.sqrt((raw.astype(np.float32) - np.float32(mn)) / np.float32(mx - mn)).astype(np.float32)
    return np.asfortranarray(Y.transpose(2, 1, 0))


@pytest.mark.parametrize("case", ["fused", "whole_page", "generic", "plain_stencil"])
def test_stencil_signs_equal_the_fp64_decisions_on_margin_inputs(case):
    """On margin inputs (every sign decided by >= 0.01) the sign bits the GPU
    exports (fpm_opts.stencil_signs) are exactly the FP64 oracle's own
    decisions, pixel for pixel (the export's index map), for the fused
    stencil kernel (n = 64 and a whole-page workgroup at n = 128), the generic
    path (n = 48) and the plain stencil kernel (4096 x 32 pupil: rows beyond
    the fused kernel's 2048); the gradient matches FP64 within 1e-5."""
    if case == "fused":
        O, P, Y, led = margin_problem(64, 4, 5)
        r = 4
    elif case == "whole_page":
        O, P, Y, led = margin_problem(128, 2, 3)
        r = 2
    elif case == "generic":
        O, P, Y, led = margin_problem(48, 4, 5)
        r = 4
    else:
        O, P, Y, led, _ = rect_margin_problem(4096, 32, 2, 6)
        r = 2
    gO, gP, _, sg = fpm_amd.fused_gradient(O, P, Y, led, r, signs=True)
    GO, GP, _, own, flips = oracle.gradient_signs(O, P, Y, led, r, signs=sg)
    print(f"[signs] {case}: flips {flips} of {2 * sg.size}; G_O {rel_l2(gO, GO):.2e} G_P {rel_l2(gP, GP):.2e}",
          flush=True)
    assert flips == 0 and np.array_equal(sg, own)
    assert rel_l2(gO, GO) <= 1e-5 and rel_l2(gP, GP) <= 1e-5


def _sign_synchronised(n, r, B, label, cap_unsync):
    """The fused gradient of the first minibatch (B images) of the n^2 x 361
    -> (n r)^2 synthetic workload at the initial object and pupil, from both
    FFT engines with their sign decisions; the FP64 oracle with each engine's
    decisions imposed, its own FP64 gradient and the FP32 oracle's."""
    ds = synth.make_device(n, r)
    N = n * r
    gpu = {}
    for name, eng in (("fp32", fpm_amd.FFT_F32), ("split", fpm_amd.FFT_SPLIT_F16)):
        GO, GP = torch.empty_like(ds["O0_dev"]), torch.empty_like(ds["P0_dev"])
        sg = torch.empty((B, n, n), dtype=torch.uint8, device=GO.device)
        torch.cuda.synchronize()
        fpm_amd.gradient_device(ds["O0_dev"].data_ptr(), ds["P0_dev"].data_ptr(), ds["raw_dev"].data_ptr(), True,
                                ds["led_dev"].data_ptr(), (N, N), (n, n), B, r, GO.data_ptr(), GP.data_ptr(),
                                engine=eng, signs_ptr=sg.data_ptr())
        torch.cuda.synchronize()
        gpu[name] = (_host(GO), _host(GP), np.asfortranarray(sg.cpu().numpy().transpose(2, 1, 0)))
        del GO, GP, sg
    Y = _prepared(ds, B)
    O0, P0 = _host(ds["O0_dev"]), _host(ds["P0_dev"])
    led = np.asfortranarray(ds["led"][:, :B])
    del ds
    torch.cuda.empty_cache()
    GO64, GP64, _, margin = oracle.gradient(O0, P0, Y, led, r, prec="f64")
    GO32, GP32, _, _ = oracle.gradient(O0, P0, Y, led, r, prec="f32")
    ref = {"G_O": _rel(GO32, GO64), "G_P": _rel(GP32, GP64)}
    print(f"[{label}] FP32 oracle vs FP64: G_O {ref['G_O']:.2e} G_P {ref['G_P']:.2e} (smallest sign margin "
          f"{margin:.1e})", flush=True)
    decisions = 2 * Y.size
    for name, (go, gp, sg) in gpu.items():
        SO, SP, _, _, flips = oracle.gradient_signs(O0, P0, Y, led, r, signs=sg, prec="f64")
        FO, FP, _, _, _ = oracle.gradient_signs(O0, P0, Y, led, r, signs=sg, prec="f32")
        es = {"G_O": _rel(go, SO), "G_P": _rel(gp, SP)}
        floor = {"G_O": _rel(FO, SO), "G_P": _rel(FP, SP)}
        e = {"G_O": _rel(go, GO64), "G_P": _rel(gp, GP64)}
        print(f"[{label}] {name} engine: sign-synchronised vs FP64 G_O {es['G_O']:.2e} G_P {es['G_P']:.2e} "
              f"(FP32 oracle with the same signs: G_O {floor['G_O']:.2e} G_P {floor['G_P']:.2e}); {flips} of "
              f"{decisions} decisions ({flips / decisions:.1e}) differ from FP64's; unsynchronised "
              f"G_O {e['G_O']:.2e} G_P {e['G_P']:.2e}", flush=True)
        assert es["G_O"] <= 1e-5, (name, es)
        # G_P = sum over the batch of R conj(S) cancels strongly: any FP32
        # computation of R and S (transforms) lands ~1e-5..3e-5 from FP64
        # there; the split engine's 22-bit products add up to ~2x more
        f = 1.5 if name == "fp32" else 3.0
        assert es["G_P"] <= max(1e-5, f * floor["G_P"]), (name, es, floor)
        assert flips <= 1e-4 * decisions, (name, flips)
        # Unsynchronised, G_O's distance comes from many near-tie flips and is
        # stable (within 1.25x of the FP32 oracle's); G_P's comes from the few
        # flips that carry much of its norm, so which decisions flip sets it
        # (split engine at 512 -> 4096: 1e-4 .. 7e-4 between builds that differ
        # only in where products are fused): the cap only there.  The
        # arithmetic is bounded by the synchronised bars above.
        _close_enough(name, {"G_O": e["G_O"]}, ref, {"G_O": 1.25}, cap_unsync)
        assert e["G_P"] <= cap_unsync, (name, e)


@pytest.mark.timeout(900)
def test_full_scale_gradient_sign_synchronised_within_1e5_of_fp64():
    """2048^2 x 361 -> 16384^2, first minibatch (26 images): each engine's
    G_O within 1e-5 of the FP64 oracle that takes the engine's own sign
    decisions (fewer than 1e-4 of the decisions differ from FP64's), G_P
    within 1e-5 or 1.5x the FP32 oracle's distance with the same signs
    (G_P = sum_l R_l conj(S_l) cancels: measured 1.2e-5 on the GPU, 2.2e-5
    for the FP32 restatement at 512 -> 4096, with FP32 or FP64 summation);
    unsynchronised, G_O within 1.25x of the FP32 oracle's distance to FP64
    (its ~1.3e-3 comes from many near-tie flips and is stable) and both
    within the cap, 2e-3 (G_P's distance comes from a few flips that carry
    much of its norm, so which decisions flip sets it)."""
    _sign_synchronised(2048, 8, 26, "fp64-full", 2e-3)


@pytest.mark.timeout(600)
def test_512_to_4096_gradient_sign_synchronised_within_1e5_of_fp64():
    """The same at 512^2 x 361 -> 4096^2 (first minibatch of 26)."""
    _sign_synchronised(512, 8, 26, "fp64-4096", 5e-3)


def test_short_solve_from_an_aberrated_pupil_as_close_to_fp64_as_the_fp32_oracle():
    """256^2 x 361 -> 1024^2, batch 26, 4 epochs of the reference's schedule
    from the plain init object and the true (aberrated) pupil, the start a
    warm-started solve has: O and P of both engines within 1.25x (FP32
    engine) / 1.5x (split engine) of the FP32 oracle's distance to FP64,
    which is about 1.07e-3 for both; cap 5e-3."""
    n, r, ep, B = 256, 4, 4, 26
    opt = (5e-4, 0.999, 1e-4)
    _, _, ds = warmstart.make_sample(n, r, 77000, pupil_sigma=warmstart.PUPIL_SIGMA)
    res = {}
    for name, eng in (("fp32", fpm_amd.FFT_F32), ("split", fpm_amd.FFT_SPLIT_F16)):
        _, _, Od, Pd = warmstart._timed_solve(ds, ds["O0_dev"], ep, B, opt, eng, halve_from=13, P0=ds["P_true_dev"])
        res[name] = (_host(Od), _host(Pd))
    O0, P0 = _host(ds["O0_dev"]), _host(ds["P_true_dev"])
    Y, led, A = _prepared(ds, ds["L"]), ds["led"], ds["A"]
    O64, P64, _ = oracle.solve(O0, P0, Y, led, A, r, ep, B, opt, prec="f64", halve_from=13)
    O32, P32, _ = oracle.solve(O0, P0, Y, led, A, r, ep, B, opt, prec="f32", halve_from=13)
    ref = {"O": _rel(O32, O64), "P": _rel(P32, P64)}
    print(f"[fp64-solve] FP32 oracle vs FP64: O {ref['O']:.2e} P {ref['P']:.2e}", flush=True)
    for name, (O, P) in res.items():
        e = {"O": _rel(O, O64), "P": _rel(P, P64)}
        print(f"[fp64-solve] {name} engine vs FP64: O {e['O']:.2e} P {e['P']:.2e}; vs FP32 oracle "
              f"O {_rel(O, O32):.2e} P {_rel(P, P32):.2e}", flush=True)
        f = 1.25 if name == "fp32" else 1.5
        _close_enough(name, e, ref, {"O": f, "P": f}, 5e-3)


@pytest.mark.timeout(900)
def test_warm_start_ablation_starts_as_close_to_fp64_as_the_fp32_oracle():
    """256^2 x 361 -> 1024^2 (the warm-start configuration), one held-out
    object: each start of the ablation -- plain object + ideal pupil (20
    epochs), network object + ideal pupil, plain object + parallax pupil,
    network object + parallax pupil (each stopped at the plain solve's final
    loss) -- against the FP64 oracle run from the same start for the same
    epochs: O and P no further from FP64 than the FP32 oracle (1.5x + 1e-4).
    The network here is trained briefly (256 objects, 2 epochs): the check
    is the solve's accuracy from each start, the time-to-loss of each start
    with the fully trained network is the benchmark's ablation."""
    n, r, B = 256, 4, 26
    opt = (5e-4, 0.999, 1e-4)
    F, T, Z = warmstart.make_training_set(256, n, r, seed0=1000, pupil_sigma=warmstart.PUPIL_SIGMA)
    net = warmstart.WarmStartNet(n, r, levels=4, base=32, max_batch=8, seed=0)
    warmstart.train(net, F, T, epochs=2, batch=8, lr=5e-4, lr_final=5e-5)
    _, _, nai = fpm_amd.led_offsets(n, r, wavelength_um=synth.WAVELENGTH_UM, na=synth.NA, pixel_um=synth.PIXEL_UM,
                                    led_pitch_um=synth.LED_PITCH_UM, led_height_um=synth.LED_HEIGHT_UM)
    pupil = warmstart.ParallaxPupil(warmstart.brightfield_planes(nai, synth.NA))
    pupil.fit(F, Z)
    del F, T, Z
    out, res, ds = warmstart.ablation(net, pupil, n, r, 77000, repeats=1)
    net.close()
    Y, led, A = _prepared(ds, ds["L"]), np.asfortranarray(ds["led"]), np.asfortranarray(ds["A"])
    for name, (O0, P0, ep, Of, Pf) in res.items():
        O0h, P0h = _host(O0), _host(P0)
        O64, P64, _ = oracle.solve(O0h, P0h, Y, led, A, r, ep, B, opt, prec="f64")
        O32, P32, _ = oracle.solve(O0h, P0h, Y, led, A, r, ep, B, opt, prec="f32")
        e = {"O": _rel(_host(Of), O64), "P": _rel(_host(Pf), P64)}
        ref = {"O": _rel(O32, O64), "P": _rel(P32, P64)}
        info = out["starts"].get(name, {})
        print(f"[ablation] {name}: {ep} epochs (reached {info.get('reached', True)}); GPU vs FP64 O {e['O']:.2e} "
              f"P {e['P']:.2e}; FP32 oracle vs FP64 O {ref['O']:.2e} P {ref['P']:.2e}", flush=True)
        _close_enough(name, e, ref, {"O": 1.5, "P": 1.5}, 2e-2)
