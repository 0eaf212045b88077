"""Sides that are not powers of two (even 2^a 3^b: 48 px -> 192, 192 px ->
768 and the like), which run the generic path (mixed-radix transforms,
unfused kernels in natural order): the transform against numpy, the fused
gradient against the FP64 oracle at <= 1e-5 (inputs whose loss-stencil signs
all have a margin), one solve step against FP64 at <= 1e-6, a 20-epoch solve
on real images as close to FP64 as the FP32 oracle, the image model against
numpy, the graph
replay against eager launches, and tiles of 32/48/64 px (48 generic) against
their own solves, bit for bit."""
import numpy as np
import pytest

import fpm_amd
import oracle
from fpm_amd import synth
from fpm_amd import tiles as ft
from helpers import margin_problem, rect_margin_problem, rel_l2, small_leds

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.mark.parametrize("rows,cols,batch", [(192, 192, 2), (768, 768, 1), (48, 96, 3), (192, 256, 1), (24, 18, 4),
                                             (3072, 48, 1)])
def test_generic_fft_matches_numpy(rows, cols, batch):
    rng = np.random.default_rng(rows + cols)
    x = np.asfortranarray((rng.standard_normal((rows, cols, batch)) + 1j * rng.standard_normal((rows, cols, batch)))
                          .astype(np.complex64))
    ref = np.fft.fft2(x.astype(np.complex128), axes=(0, 1))
    got = np.stack([fpm_amd.fft2(x[:, :, k]) for k in range(batch)], axis=2)
    assert rel_l2(got, ref) < 2e-6
    inv = np.fft.ifft2(x.astype(np.complex128), axes=(0, 1)) * rows * cols
    goti = np.stack([fpm_amd.fft2(x[:, :, k], inverse=True) for k in range(batch)], axis=2)
    assert rel_l2(goti, inv) < 2e-6


@pytest.mark.parametrize("rows,cols,batch", [(48, 96, 3), (96, 48, 2), (384, 24, 1), (1536, 20, 1), (6144, 18, 1),
                                             (12288, 16, 1), (20, 768, 2), (26, 384, 1), (30, 1536, 1), (50, 192, 1)])
def test_3x2k_line_kernel_matches_numpy(rows, cols, batch):
    """Lines of 3 x 2^k points run a radix-3 step into three power-of-two
    transforms (fft3m_kernel): every M from 16 to 4096 along the contiguous
    index (48 .. 12288 rows), M <= 256 along the strided index (48 .. 768
    columns; 1536 columns take the generic kernel), workgroups holding fewer
    lines than their capacity (24 lines of 48, 50 rows of 192 ...), both
    directions, against numpy in FP64."""
    rng = np.random.default_rng(rows * 7 + cols)
    x = np.asfortranarray((rng.standard_normal((rows, cols, batch)) + 1j * rng.standard_normal((rows, cols, batch)))
                          .astype(np.complex64))
    ref = np.fft.fft2(x.astype(np.complex128), axes=(0, 1))
    got = np.stack([fpm_amd.fft2(x[:, :, k]) for k in range(batch)], axis=2)
    assert rel_l2(got, ref) < 2e-6
    inv = np.fft.ifft2(x.astype(np.complex128), axes=(0, 1)) * rows * cols
    goti = np.stack([fpm_amd.fft2(x[:, :, k], inverse=True) for k in range(batch)], axis=2)
    assert rel_l2(goti, inv) < 2e-6


@pytest.mark.parametrize("n,r,grid", [(48, 4, 5), (24, 2, 3), (96, 2, 3)])
def test_generic_gradient_matches_fp64_oracle(n, r, grid):
    O, P, Y, led = margin_problem(n, r, grid)
    GO, GP, loss, margin = oracle.gradient(O, P, Y, led, r)
    assert margin > 5e-3
    gO, gP, gl = fpm_amd.fused_gradient(O, P, Y, led, r)
    print(f"[generic] gradient n={n} r={r}: G_O {rel_l2(gO, GO):.2e} G_P {rel_l2(gP, GP):.2e}", flush=True)
    assert rel_l2(gO, GO) <= 1e-5
    assert rel_l2(gP, GP) <= 1e-5
    assert abs(gl - loss) <= 1e-5 * loss


def test_generic_solve_one_step_matches_fp64_oracle():
    n, r = 48, 4
    O, P, Y, led = margin_problem(n, r, 5)
    A, _ = synth.pupil(n)
    L = led.shape[1]
    opt = (2e-3, 0.999, 1e-4)
    O1, P1 = fpm_amd.cuFPM_pure_v3(O, P, Y, led, A, r, 1, L, opt)
    Oo, Po, _ = oracle.solve(O, P, Y, led, A, r, 1, L, opt, prec="f64")
    assert rel_l2(O1, Oo) < 1e-6
    assert rel_l2(P1, Po) < 1e-6


def test_generic_solve_epochs_on_margin_inputs_match_fp64():
    """3 epochs over 25 LEDs in batches of 10 (ragged last batch), small
    steps that keep the margin inputs' sign decisions: O, P and every epoch's
    loss vs the FP64 oracle (the moments carry across steps and epochs)."""
    n, r = 48, 4
    O, P, Y, led = margin_problem(n, r, 5)
    A, _ = synth.pupil(n)
    opt = (2e-4, 0.999, 1e-4)
    O1, P1, lg = fpm_amd.cuFPM_pure_v3(O, P, Y, led, A, r, 3, 10, opt, loss=True)
    Oo, Po, lo = oracle.solve(O, P, Y, led, A, r, 3, 10, opt, prec="f64")
    print(f"[generic] 3 epochs margin: O {rel_l2(O1, Oo):.2e} P {rel_l2(P1, Po):.2e}", flush=True)
    assert rel_l2(O1, Oo) < 1e-5
    assert rel_l2(P1, Po) < 1e-5
    np.testing.assert_allclose(lg, lo, rtol=1e-5)


def _dataset(n, r, grid):
    led = small_leds(n, r, grid)
    A, Ptrue = synth.pupil(n)
    amp = synth.object_numpy(n * r, r, 0.3, 11)
    model = synth.image_model_numpy(np.asfortranarray(amp.astype(np.complex128)), Ptrue.astype(np.complex128), led, r)
    Y = np.asfortranarray(model.astype(np.float32))
    O0 = np.asfortranarray(synth.bilinear_upsample(Y[:, :, 0].astype(np.float64), r).astype(np.complex64))
    return O0, np.asfortranarray(A.astype(np.complex64)), Y, led, A


def test_generic_solve_20_epochs_as_close_to_fp64_as_the_fp32_oracle():
    """20 epochs on real (noise-free model) images: loss-stencil signs near a
    tie flip differently in any two FP32 computations (here two different FFT
    algorithms), and the trajectories part chaotically: measured 9.6e-4 (O)
    and 1.3e-3 (P) from FP64 against the FP32 oracle's 6.0e-4 and 5.0e-4.
    Bars: below 2e-3 from FP64 and from the FP32 oracle, losses within 1e-3
    (the exact regime is pinned by the margin tests)."""
    n, r, grid = 48, 4, 7
    O0, P0, Y, led, A = _dataset(n, r, grid)
    opt = (5e-4, 0.999, 1e-4)
    O, P, lg = fpm_amd.cuFPM_pure_v3(O0, P0, Y, led, A, r, 20, 10, opt, loss=True)
    Oo, Po, lo = oracle.solve(O0, P0, Y, led, A, r, 20, 10, opt, prec="f32")
    O64, P64, _ = oracle.solve(O0, P0, Y, led, A, r, 20, 10, opt, prec="f64")
    e = {"O": rel_l2(O, O64), "P": rel_l2(P, P64)}
    ref = {"O": rel_l2(Oo, O64), "P": rel_l2(Po, P64)}
    print(f"[generic] 20 epochs vs FP64: GPU O {e['O']:.2e} P {e['P']:.2e}; FP32 oracle O {ref['O']:.2e} "
          f"P {ref['P']:.2e}; GPU vs FP32 oracle O {rel_l2(O, Oo):.2e} P {rel_l2(P, Po):.2e}", flush=True)
    for k in e:
        assert e[k] <= 2e-3, (k, e[k], ref[k])
    assert rel_l2(O, Oo) <= 2e-3 and rel_l2(P, Po) <= 2e-3
    assert np.allclose(lg, lo, rtol=1e-3)
    assert lg[-1] < lg[0]


def test_generic_graph_replay_equals_eager():
    n, r, grid = 48, 4, 7
    O0, P0, Y, led, A = _dataset(n, r, grid)
    opt = (5e-4, 0.999, 1e-4)
    a = fpm_amd.cuFPM_pure_v3(O0, P0, Y, led, A, r, 4, 10, opt, loss=True)
    b = fpm_amd.cuFPM_pure_v3(O0, P0, Y, led, A, r, 4, 10, opt, loss=True, graph=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    np.testing.assert_allclose(a[2], b[2], rtol=1e-12)


def test_generic_image_model_matches_numpy():
    n, r = 48, 4
    led = small_leds(n, r, 5)
    _, P = synth.pupil(n)
    O = np.asfortranarray(synth.object_numpy(n * r, r, 0.3, 5).astype(np.complex64))
    got = fpm_amd.forward_intensity(O, P, led, r)
    ref = synth.image_model_numpy(O.astype(np.complex128), P.astype(np.complex128), led, r)
    assert rel_l2(got, ref) < 2e-6


@pytest.mark.parametrize("launch", [ft.TILES_GROUPED, ft.TILES_PER_TILE])
def test_generic_tiles_equal_their_own_solves(launch):
    """Tiles of 32, 48 and 64 px (the 48-px ones on the generic path) in one
    call, grouped per size (the generic kernels take a tile grid row) or one
    launch sequence per tile: every tile equals its own solve."""
    tl = ft.make_tile_set(1024, (32, 48, 64), 3, 4, seed=4, led_grid=7)
    rep, _ = ft.tiles_solve(tl, 4, 2, 10, (1e-3, 0.999, 1e-4), streams=2, tile_launch=launch)
    print(f"[generic] tiles: report {rep}", flush=True)
    assert sum(1 for t in tl if t["n"] == 48) == 3
    if launch == ft.TILES_GROUPED:
        assert rep["grouped_tiles"] == len(tl) and rep["sequences"] == 3
    else:
        assert rep["grouped_tiles"] == 0 and rep["sequences"] == len(tl)
    for k, t in enumerate(tl):
        O, P = torch.empty_like(t["O0"]), torch.empty_like(t["P0"])
        N, n = t["O0"].shape[0], t["P0"].shape[0]
        torch.cuda.synchronize()
        fpm_amd.solve_device(t["O0"].data_ptr(), t["P0"].data_ptr(), t["Y"].data_ptr(), True, t["led"].data_ptr(),
                             t["A"].data_ptr(), (N, N), (n, n), t["Y"].shape[0], 4, 2, 10, (1e-3, 0.999, 1e-4),
                             O.data_ptr(), P.data_ptr())
        assert np.array_equal(t["O"].cpu().numpy(), O.cpu().numpy()), k
        assert np.array_equal(t["P"].cpu().numpy(), P.cpu().numpy()), k


# ------------------------------------------------- any even side up to 16384
@pytest.mark.parametrize("rows,cols,batch", [(200, 40, 2), (20, 98, 1), (134, 22, 3), (254, 16, 1),
                                             (12288, 16, 1), (16, 12288, 1), (1600, 200, 1), (16, 8186, 1)])
def test_generic_fft_radix_5_7_and_prime_factors_match_numpy(rows, cols, batch):
    """Lengths with factors 5 and 7 (hard-coded butterflies), primes 67, 127
    and 4093 (direct sums: 134 = 2 67, 254 = 2 127, 8186 = 2 4093) and lines
    of 12288 points (one line per workgroup, 96 KiB of LDS)."""
    rng = np.random.default_rng(rows * 7 + cols)
    x = np.asfortranarray((rng.standard_normal((rows, cols, batch)) + 1j * rng.standard_normal((rows, cols, batch)))
                          .astype(np.complex64))
    ref = np.fft.fft2(x.astype(np.complex128), axes=(0, 1))
    got = np.stack([fpm_amd.fft2(x[:, :, k]) for k in range(batch)], axis=2)
    inv = np.fft.ifft2(x.astype(np.complex128), axes=(0, 1)) * rows * cols
    goti = np.stack([fpm_amd.fft2(x[:, :, k], inverse=True) for k in range(batch)], axis=2)
    print(f"[generic-fft] {rows} x {cols}: forward {rel_l2(got, ref):.2e} inverse {rel_l2(goti, inv):.2e}", flush=True)
    # a direct p-point sum in FP32 adds ~sqrt(p) eps per pass
    tol = 2e-6 if max(rows, cols) <= 12288 and cols != 8186 else 2e-5
    assert rel_l2(got, ref) < tol and rel_l2(goti, inv) < tol


@pytest.mark.timeout(900)
@pytest.mark.parametrize("n,r,grid", [(200, 8, 3), (134, 2, 3), (1536, 8, 3)])
def test_generic_any_even_side_gradient_and_step_match_fp64(n, r, grid):
    """200 -> 1600 (2^6 5^2), 134 -> 268 (4 67: a prime factor summed
    directly) and 1536 -> 12288 (3 4096: 12288-point lines) on margin inputs:
    the fused gradient within 1e-5 of FP64 and one solve step within 1e-6.
    (The 12288 object is a band-limited random field with LED windows near
    the centre: the synthetic specimen's 20 Gaussians cost minutes there.)"""
    if n == 1536:
        O, P, Y, led, _ = rect_margin_problem(n, n, r, 6)
    else:
        O, P, Y, led = margin_problem(n, r, grid)
    GO, GP, loss, margin = oracle.gradient(O, P, Y, led, r)
    assert margin > 5e-3
    gO, gP, gl = fpm_amd.fused_gradient(O, P, Y, led, r)
    print(f"[generic-any] {n} -> {n * r}: gradient G_O {rel_l2(gO, GO):.2e} G_P {rel_l2(gP, GP):.2e}", flush=True)
    assert rel_l2(gO, GO) <= 1e-5 and rel_l2(gP, GP) <= 1e-5
    assert abs(gl - loss) <= 1e-5 * loss
    A, _ = synth.pupil(n)
    L = led.shape[1]
    opt = (2e-3, 0.999, 1e-4)
    O1, P1 = fpm_amd.cuFPM_pure_v3(O, P, Y, led, A, r, 1, L, opt)
    Oo, Po, _ = oracle.solve(O, P, Y, led, A, r, 1, L, opt, prec="f64")
    print(f"[generic-any] {n} -> {n * r}: one step O {rel_l2(O1, Oo):.2e} P {rel_l2(P1, Po):.2e}", flush=True)
    assert rel_l2(O1, Oo) < 1e-6 and rel_l2(P1, Po) < 1e-6


def test_generic_launch_shapes_do_not_change_results():
    """A solve's set-up times candidate launch shapes of its transform passes
    and keeps the fastest per (length, line kind, source); the shape only
    spreads lines over workgroups and threads, so a transform of lengths no
    other test uses gives the same bits before (default shapes) and after a
    solve of that geometry (tuned shapes)."""
    n, r = 210, 4  # 2 3 5 7 -> 840 = 2^3 3 5 7: lengths of no other test
    rng = np.random.default_rng(840)
    x = np.asfortranarray((rng.standard_normal((n * r, n * r, 2)) +
                           1j * rng.standard_normal((n * r, n * r, 2))).astype(np.complex64))
    before = fpm_amd.fft2(x)
    O, P, Y, led = margin_problem(n, r, 3)
    A, _ = synth.pupil(n)
    fpm_amd.cuFPM_pure_v3(O, P, Y, led, A, r, 1, led.shape[1], (1e-3, 0.999, 1e-4))
    after = fpm_amd.fft2(x)
    assert np.array_equal(np.ascontiguousarray(before).view(np.uint32), np.ascontiguousarray(after).view(np.uint32))
    inv = fpm_amd.fft2(after, inverse=True) / (n * r) ** 2
    assert rel_l2(inv, x) < 3e-6
