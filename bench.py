"""Benchmark of the fused FPM solve on one MI355X.

Workload: the reference's headline case -- 361 uint16 images of 2048 x 2048,
upsampling 8 -> a 16384 x 16384 complex64 object, batch 26, FP32 -- on
synthetic data (fpm_amd.synth.make_device; the reference's data are external).
A "step" is one epoch (one pass over all 361 images: 14 minibatch steps).
The timed region is one fpm_pure_v3 call of K epochs on device-resident
inputs, bracketed by device synchronisation, exactly as the reference's
mex call is timed around cuFPM_pure_v3 (README.md:36 quotes the whole call).

Also reported:
  stages     per-stage device time (HIP events on the solve's stream, inside
             the timed region), GB/s against the dense byte model of the
             unfused steps (per step 104 N^2, per LED 138 n^2 with uint16 Y)
             and, for the fused kernels, against the bytes they must move;
  roofline   the dominant kernel: the bytes it must move per launch over its
             event-timed duration, against MI355X's 8 TB/s HBM peak and a
             plain device copy timed here, and its measured HBM traffic
             (rocprofv3 counters, copy-calibrated, from profiles/);
  batch_autotune the opt-in minibatch autotuner (fpm_amd.tune): seconds to
             the batch-26 solve's final loss for each candidate size;
  fft_reference the library's 2D DFT beside rocFFT (torch.fft) on the
             16384^2 spectrum and 26 pages of 2048^2, same buffers and stream;
  cpu_baseline the FP32 CPU restatement (oracle, OpenMP) on the host cores:
             one epoch of the same workload, plus the 128^2 -> 512^2 case over
             20 epochs and the 512^2 -> 4096^2 case over 2 epochs.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fused-fourier-ptycho_amd"))

METRIC = "s/epoch 2048²×361→16384²; HBM GB/s vs peak; time-to-loss vs plain @rel-L2 1e-3"
HBM_PEAK_GBPS = 8000.0
STAGE_NAMES = ["spectrum_init", "crop", "ifft_n_cols", "ifft_n_contig", "stencil", "fft_n_contig", "fft_n_cols",
               "zero", "stitch_owner", "backproj", "ifft_N_cols", "fused_update", "fft_N_cols"]


def stage_alg_bytes(n, N, b_avg, y_bytes, stitch):
    """Dense algorithmic bytes of ONE execution of each stage.

    Large-array terms per step (c64 = 8 B, f32 = 4 B): DFT(O) 32 N^2 and
    IDFT(G) 32 N^2 as two passes of (8 read + 8 write), RMSprop with gradient
    zeroing 40 N^2 (read G, O, m; write O, m, G=0).  Per-LED terms: crop 24 n^2
    (read S, write sub + latent), each batched transform 32 n^2, stencil
    16 n^2 + Y, back-projection 32 n^2 (read R, sub; read-modify-write G)."""
    NN, nn = float(N) * N, float(n) * n
    lo = b_avg * nn
    d = {
        "spectrum_init": 32 * NN,
        "crop": 24 * lo,
        "ifft_n_cols": 16 * lo,
        "ifft_n_contig": 16 * lo,
        "stencil": (16 + y_bytes) * lo,
        "fft_n_contig": 16 * lo,
        "fft_n_cols": 16 * lo,
        "zero": 8 * NN,
        # owner-computes: reads R once, writes every spectrum element (the
        # clearing it replaces); the back-projection then reads R and sub
        "stitch_owner": 8 * lo + 8 * NN,
        "backproj": (32 if stitch == "atomic" else 16) * lo,
        "ifft_N_cols": 16 * NN,
        # contiguous IDFT pass 16 + RMSprop without the zeroing 32 + contiguous DFT pass 16
        "fused_update": 64 * NN,
        "fft_N_cols": 16 * NN,
    }
    return d


def window_rows_per_step(led, n, N, batch):
    """Mean number of object-spectrum rows that the LED windows of a
    minibatch cover (union over its LEDs of rows (kt - 1 + N/2 + i) mod N,
    i < n, cuFPM_v3/cuda/kernels.cuh:43-47): outside them the step's
    spectrum gradient is zero by construction, so a kernel that reads the
    gradient needs only these rows."""
    import numpy as np

    L = led.shape[1]
    rows = []
    for c in range(0, L, batch):
        hit = np.zeros(N, dtype=bool)
        for kt in led[1, c:c + batch]:
            hit[(int(kt) - 1 + N // 2 + np.arange(n)) % N] = True
        rows.append(int(hit.sum()))
    return float(np.mean(rows))


def stage_necessary_bytes(n, N, b_avg, y_bytes, win_rows):
    """HBM bytes the fused kernels of a stage must move at least (their own
    inputs and outputs once each; no credit for the stages they absorb):
    fused_update reads O and m_O and the gradient rows inside the step's
    windows and writes O, m_O and the next spectrum's rows inside the next
    step's windows: 24 N^2 + 8 N (rows + next rows), with the mean window
    rows for both; the fused stencil reads the inverse-transformed pages'
    inputs and Y and writes the pages: (8 + 8 + y) b n^2."""
    NN, nn = float(N) * N, float(n) * n
    return {"fused_update": 24 * NN + 16.0 * N * win_rows, "stencil": (16 + y_bytes) * b_avg * nn}


# stage -> kernel whose per-dispatch HBM traffic stands for it in the PMC summary
STAGE_KERNEL = {"fused_update": "fused_update16k_kernel<true>", "stencil": "stencil_half_kernel"}
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "hbm_traffic_fpm_solve_16384_b26.md")


def copy_bandwidth_gbps(nbytes=1 << 30, reps=5):
    """HBM bandwidth of a plain device-to-device copy of nbytes (read + write
    counted), best of `reps`, timed with HIP events on the current stream."""
    import torch

    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    dst.copy_(src)
    best = float("inf")
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dst.copy_(src)
        e1.record()
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1) / 1e3)
    del src, dst
    torch.cuda.empty_cache()
    return 2.0 * nbytes / best / 1e9


def measured_traffic(stage, n, N):
    """Read + write bytes per dispatch of the stage's kernel from the committed
    rocprofv3 PMC summary (FETCH_SIZE / WRITE_SIZE, copy-calibrated), or None."""
    kern = STAGE_KERNEL.get(stage)
    if kern is None or (n, N) != (2048, 16384) or not os.path.exists(TRAFFIC_SUMMARY):
        return None
    for line in open(TRAFFIC_SUMMARY):
        cells = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
        if len(cells) == 4 and cells[0].endswith(kern):
            return (float(cells[2]) + float(cells[3])) * 1e9
    return None


def run_fft_reference(stream, reps=5):
    """The library's unnormalised 2D DFT beside rocFFT (torch.fft) on the same
    device buffers and stream: the object spectrum (16384^2) and one
    minibatch of per-LED pages (26 x 2048^2), forward, best of `reps`, HIP
    events on the stream both run on; also their relative L2 distance."""
    import torch

    import fpm_amd

    res = {}
    for rows, cols, batch in ((16384, 16384, 1), (2048, 2048, 26)):
        g = torch.Generator(device="cuda:0").manual_seed(7)
        x = torch.complex(torch.randn((batch, cols, rows), generator=g, device="cuda:0"),
                          torch.randn((batch, cols, rows), generator=g, device="cuda:0"))
        y = torch.empty_like(x)

        def ours():
            fpm_amd._check(fpm_amd.lib().fpm_fft2_engine(x.data_ptr(), y.data_ptr(), rows, cols, batch, -1,
                                                         fpm_amd.FFT_F32, 1, stream.cuda_stream))

        def timed(fn):
            fn()
            best = float("inf")
            for _ in range(reps):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                fn()
                e1.record(stream)
                e1.synchronize()
                best = min(best, e0.elapsed_time(e1))
            return best

        with torch.cuda.stream(stream):
            ms_ours = timed(ours)
            ref = torch.fft.fft2(x)
            ms_roc = timed(lambda: torch.fft.fft2(x))
            rel = ((y - ref).abs().double().pow(2).sum() / ref.abs().double().pow(2).sum()).sqrt().item()
        moved = 2.0 * x.numel() * 8  # one read + one write of the pages
        res[f"{batch}x{rows}x{cols}"] = {"ms": round(ms_ours, 4), "rocfft_ms": round(ms_roc, 4),
                                         "GBps_read_write_once": round(moved / 1e6 / ms_ours, 1),
                                         "rocfft_GBps_read_write_once": round(moved / 1e6 / ms_roc, 1),
                                         "rel_l2_vs_rocfft": rel}
        del x, y, ref
        torch.cuda.empty_cache()
    return res


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _host_problem(ds, count=None):
    """(O0, P0, Y, led, A) on the host from a synth.make_device dict: Y the
    first `count` images (all if None) prepared as the solve prepares them."""
    import numpy as np

    L = ds["L"] if count is None else count
    raw = ds["raw_dev"][:L].cpu().numpy()  # (L, n, n) row-major = column-major pages
    mn, mx = ds["raw_min"], ds["raw_max"]
    Y = np.sqrt((raw.astype(np.float32) - np.float32(mn)) / np.float32(mx - mn)).astype(np.float32)
    del raw
    Y = np.asfortranarray(Y.transpose(2, 1, 0))
    O0 = np.asfortranarray(ds["O0_dev"].cpu().numpy().T)
    P0 = np.asfortranarray(ds["P0_dev"].cpu().numpy().T)
    return O0, P0, Y, np.asfortranarray(ds["led"][:, :L]), np.asfortranarray(ds["A"])


def run_cpu_baseline(ds, args):
    """The FP32 CPU restatement (oracle, OpenMP over the host cores), timed
    as the plan of the CPU baseline asks: one full epoch of the benchmark's
    workload (all images, every minibatch step: the reported value), the
    128^2 x 361 -> 512^2 case over its 20 epochs and the 512^2 x 361 ->
    4096^2 case over 2 epochs (same generator, batch, optimizer)."""
    import torch

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    from fpm_amd import synth

    oracle.lib()
    opt = (args.lr, 0.999, 1e-4)
    B = args.batch
    n, N = ds["n"], ds["N"]
    O0, P0, Y, led, A = _host_problem(ds)
    t0 = time.perf_counter()
    oracle.solve(O0, P0, Y, led, A, ds["pratio"], 1, B, opt, prec="f32")
    t_epoch = time.perf_counter() - t0
    del O0, P0, Y
    print(f"[bench] cpu baseline: one epoch of {n}^2 -> {N}^2 in {t_epoch:.1f} s", file=sys.stderr, flush=True)
    plan = {}
    for (sn, sr, epochs) in ((128, 4, 20), (512, 8, 2)):
        small = synth.make_device(sn, sr, device="cuda:0")
        hp = _host_problem(small)
        del small
        torch.cuda.empty_cache()
        t0 = time.perf_counter()
        oracle.solve(*hp[:4], hp[4], sr, epochs, B, opt, prec="f32")
        dt = time.perf_counter() - t0
        plan[f"{sn}^2x361->{sn * sr}^2"] = {"epochs": epochs, "s_total": round(dt, 3),
                                            "s_per_epoch": round(dt / epochs, 4)}
        print(f"[bench] cpu baseline: {sn}^2 -> {sn * sr}^2, {epochs} epochs in {dt:.1f} s", file=sys.stderr,
              flush=True)
    return {"value": t_epoch, "unit": "s/epoch", "cores": oracle.num_threads(), "kind": "port",
            "cpu_model": _cpu_model(),
            "sample": f"one full epoch ({(ds['L'] + B - 1) // B} minibatch steps over all {ds['L']} images) of the "
                      f"{n}^2 x {ds['L']} -> {N}^2 workload, FP32 C restatement of the reference's solve, OpenMP "
                      f"on {oracle.num_threads()} host threads",
            "plan": plan}


TTL_TEST_SEEDS = tuple(range(77000, 77008))


def run_time_to_loss(args):
    """Plain solve vs learned warm start + solve on the warm-start workload
    (361 uint16 images of 256 x 256 -> 1024 x 1024, batch 26, 20 epochs,
    objects imaged through random pupils: 5 Zernike coefficients N(0, 0.4)
    rad).  The U-Net (4 levels, base 32) is trained here on synthetic
    objects through the forward model, the parallax pupil map is fitted on
    the same training set, and on held-out objects the warm-started solve
    runs until it reaches the plain 20-epoch solve's final loss."""
    import numpy as np
    import torch

    import fpm_amd
    from fpm_amd import synth, warmstart

    n, r = 256, 4
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.perf_counter()
    F, T, Z = warmstart.make_training_set(args.ttl_samples, n, r, seed0=1000, pupil_sigma=warmstart.PUPIL_SIGMA)
    t_gen = time.perf_counter() - t0
    net = warmstart.WarmStartNet(n, r, levels=4, base=32, max_batch=8, seed=0)
    t0 = time.perf_counter()
    hist = warmstart.train(net, F, T, epochs=args.ttl_epochs, batch=8, lr=5e-4, lr_final=5e-5, log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    _, _, nai = fpm_amd.led_offsets(n, r, wavelength_um=synth.WAVELENGTH_UM, na=synth.NA, pixel_um=synth.PIXEL_UM,
                                    led_pitch_um=synth.LED_PITCH_UM, led_height_um=synth.LED_HEIGHT_UM)
    pupil = warmstart.ParallaxPupil(warmstart.brightfield_planes(nai, synth.NA))
    t0 = time.perf_counter()
    fit_rms = pupil.fit(F, Z)
    torch.cuda.synchronize()
    t_fit = time.perf_counter() - t0
    del F, T, Z
    torch.cuda.empty_cache()
    res = {}
    for name, eng in (("fp32", fpm_amd.FFT_F32), ("split", fpm_amd.FFT_SPLIT_F16)):
        per = []
        for seed in TTL_TEST_SEEDS:
            rr = warmstart.time_to_loss(net, n, r, seed, warm_engine=eng, pupil_model=pupil)
            rr.pop("loss_per_epoch")
            per.append(rr)
            log(f"[bench] time to loss ({name} warm solve, object {seed}): ratio {rr['ratio_vs_plain']:.3f}, "
                f"{rr['warm_epochs_run']} epochs, reached {rr['warm_reaches_plain_loss']}")
        ratios = [p["ratio_vs_plain"] for p in per]
        res[name] = {"mean_ratio_vs_plain": float(np.mean(ratios)), "max_ratio_vs_plain": float(np.max(ratios)),
                     "all_reached": all(p["warm_reaches_plain_loss"] for p in per),
                     "mean_epochs_saved": float(np.mean([p["epochs_saved"] for p in per])), "objects": per}
    # ablation: what the network and the parallax pupil each buy (FP32 engine)
    abl = {k: [] for k in warmstart.ABLATION_STARTS}
    for seed in TTL_TEST_SEEDS:
        rr, _, _ = warmstart.ablation(net, pupil, n, r, seed)
        for k in warmstart.ABLATION_STARTS:
            abl[k].append(dict(rr["starts"][k], plain_final_rel_l2_vs_truth=rr["plain"]["final_rel_l2_vs_truth"]))
        log(f"[bench] ablation, object {seed}: " + ", ".join(
            f"{k} {rr['starts'][k]['epochs_run']} ep ratio {rr['starts'][k]['ratio_vs_plain']}"
            for k in warmstart.ABLATION_STARTS))
    ablation = {}
    for k, per in abl.items():
        ratios = [p["ratio_vs_plain"] for p in per if p["ratio_vs_plain"] is not None]
        ablation[k] = {"reached": sum(p["reached"] for p in per), "objects": len(per),
                       "mean_ratio_vs_plain_of_reached": float(np.mean(ratios)) if ratios else None,
                       "mean_epochs_run": float(np.mean([p["epochs_run"] for p in per])),
                       "mean_final_rel_l2_vs_truth": float(np.mean([p["final_rel_l2_vs_truth"] for p in per])),
                       "mean_plain_final_rel_l2_vs_truth": float(np.mean(
                           [p["plain_final_rel_l2_vs_truth"] for p in per]))}
    net.close()
    return {"workload": f"361 uint16 images {n}x{n} -> {n * r}x{n * r}, batch 26, pupils with 5 Zernike terms "
                        f"~ N(0, {warmstart.PUPIL_SIGMA}) rad, plain 20 epochs (FP32 engine, reference schedule, "
                        f"ideal-pupil init) vs U-Net object + parallax pupil + solve stopped at the plain final "
                        f"loss, {len(TTL_TEST_SEEDS)} held-out objects",
            "training": {"objects": args.ttl_samples, "epochs": args.ttl_epochs, "batch": 8, "generate_s": t_gen,
                         "train_s": t_train, "loss_first_last": [hist[0], hist[-1]],
                         "pupil_fit_s": t_fit, "pupil_fit_rms_rad": fit_rms},
            "warm_fft_engine": res,
            "ablation_fp32_engine": ablation}


def _tile_sizes(args):
    return tuple(int(v) for v in args.tile_sizes.split(","))


def run_tiles(args):
    """Uneven field-of-view tiles: 48 tiles of 128, 192 and 256 low-res
    pixels (--tile-sizes; upsample 4) over a 1024-pixel field, 19 x 19 LEDs
    seen from each tile's centre, per-tile Zernike pupils (sigma 0.4 rad),
    batch 26, reconstructed jointly: the tiles of each size in shared
    launches (the 192-px tiles on the generic mixed-radix path),
    one hipGraph per epoch, autotuned stream count (the autotuning call's time
    includes its trial epochs and captures, so the chosen count is also timed
    fixed); the same with one launch sequence per tile; the grouped call on
    the split-FP16 engine; and the tiles solved one after another with
    fpm_pure_v3."""
    import torch

    import fpm_amd
    from fpm_amd import tiles as ft

    t0 = time.perf_counter()
    tl = ft.make_tile_set(1024, _tile_sizes(args), 7, 4, seed=11, count=48)
    t_gen = time.perf_counter() - t0
    ep, opt = args.tile_epochs, (args.lr, 0.999, 1e-4)

    def timed(**kw):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rep, _ = ft.tiles_solve(tl, 4, ep, args.batch, opt, **kw)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, rep

    # warm-up calls (first launches of every kernel shape), then the timed calls
    for launch in (ft.TILES_GROUPED, ft.TILES_PER_TILE):
        ft.tiles_solve(tl, 4, 1, args.batch, opt, streams=1, tile_launch=launch)
    ft.tiles_solve(tl, 4, 1, args.batch, opt, streams=1, engine=fpm_amd.FFT_SPLIT_F16)
    t_joint, rep = timed(streams=0)
    # steady state: the chosen stream count fixed (no tuning epochs, one capture per epoch graph)
    t_fixed, _ = timed(streams=rep["streams"])
    t_pt, rep_pt = timed(streams=0, tile_launch=ft.TILES_PER_TILE)
    t_pt_fixed, _ = timed(streams=rep_pt["streams"], tile_launch=ft.TILES_PER_TILE)
    t_split, rep_s = timed(streams=0, engine=fpm_amd.FFT_SPLIT_F16)
    t0 = time.perf_counter()
    for t in tl:
        N, n = t["O0"].shape[0], t["P0"].shape[0]
        fpm_amd.solve_device(t["O0"].data_ptr(), t["P0"].data_ptr(), t["Y"].data_ptr(), True, t["led"].data_ptr(),
                             t["A"].data_ptr(), (N, N), (n, n), t["Y"].shape[0], 4, ep, args.batch, opt,
                             t["O"].data_ptr(), t["P"].data_ptr())
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    # dense algorithmic bytes of one epoch of all tiles (the byte model of the
    # single solve: 104 N^2 per step, 138 n^2 per LED with uint16 images)
    L = int(tl[0]["Y"].shape[0])
    steps = (L + args.batch - 1) // args.batch
    epoch_bytes = sum(steps * 104.0 * t["O0"].shape[0] ** 2 + L * 138.0 * t["P0"].shape[0] ** 2 for t in tl)
    sizes = "/".join(str(v) for v in _tile_sizes(args))
    return {"workload": f"48 tiles of {sizes} px (upsample 4) over a 1024 px field, 361 LEDs per tile seen from "
                        f"its centre, per-tile pupils, batch {args.batch}, {ep} epochs",
            "grouped": {"s_per_epoch": t_joint / ep, "fixed_streams_s_per_epoch": t_fixed / ep,
                        "streams_chosen": rep["streams"], "autotune_epoch_ms": rep["candidates"],
                        "sequences": rep["sequences"], "grouped_tiles": rep["grouped_tiles"],
                        "epoch_dense_model_GBps": epoch_bytes / 1e9 / (t_fixed / ep)},
            "per_tile": {"s_per_epoch": t_pt / ep, "fixed_streams_s_per_epoch": t_pt_fixed / ep,
                         "streams_chosen": rep_pt["streams"], "autotune_epoch_ms": rep_pt["candidates"],
                         "sequences": rep_pt["sequences"]},
            "grouped_split_engine_s_per_epoch": t_split / ep, "split_streams_chosen": rep_s["streams"],
            "sequential_s_per_epoch": t_seq / ep, "speedup_grouped_vs_sequential": t_seq / t_joint,
            "speedup_grouped_vs_per_tile": t_pt / t_joint, "epoch_dense_model_GB": epoch_bytes / 1e9,
            "generate_s": t_gen,
            "note": "sides that are not powers of two run the generic path (mixed-radix transforms, unfused "
                    "kernels, grouped like the others); the split engine applies to power-of-two sides"}


def run_tiles_grouped(args):
    """The grouped call of the tiles leg alone (same tiles), fixed stream
    count, one warm-up epoch then args.tile_epochs timed."""
    import torch

    from fpm_amd import tiles as ft

    tl = ft.make_tile_set(1024, _tile_sizes(args), 7, 4, seed=11, count=48)
    ep, opt = args.tile_epochs, (args.lr, 0.999, 1e-4)
    ft.tiles_solve(tl, 4, 1, args.batch, opt, streams=args.tile_streams)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rep, _ = ft.tiles_solve(tl, 4, ep, args.batch, opt, streams=args.tile_streams)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"s_per_epoch": dt / ep, "epochs": ep, "streams": rep["streams"], "sequences": rep["sequences"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="epochs timed (one call)")
    ap.add_argument("--warmup", type=int, default=1, help="epochs of the untimed warm-up call")
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--pratio", type=int, default=8)
    ap.add_argument("--batch", type=int, default=26)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--stitch", choices=["atomic", "owner"], default="owner")
    ap.add_argument("--y", choices=["u16", "f32"], default="u16")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-sweep", default="16,26,32",
                    help="comma-separated batch sizes also timed (same epochs, s/epoch each); empty: none")
    ap.add_argument("--no-stages", action="store_true", help="time without per-stage events")
    ap.add_argument("--batch-autotune", default="16,26,32",
                    help="candidate batch sizes of the time-to-loss batch autotuner, run with the batch sweep "
                         "('' to skip)")
    ap.add_argument("--engine", choices=["f32", "split"], default="f32",
                    help="FFT radix-16 stages: FP32 butterflies or split-FP16 MFMA products")
    ap.add_argument("--graph", action="store_true",
                    help="capture each epoch as a hipGraph and replay it (per-stage events off)")
    ap.add_argument("--no-time-to-loss", action="store_true", help="skip the warm-start time-to-loss leg")
    ap.add_argument("--ttl-samples", type=int, default=4096, help="synthetic training objects of the warm start")
    ap.add_argument("--ttl-epochs", type=int, default=8, help="training epochs of the warm start")
    ap.add_argument("--no-tiles", action="store_true", help="skip the multi-tile leg")
    ap.add_argument("--no-fft-reference", action="store_true", help="skip the rocFFT (torch.fft) comparison")
    ap.add_argument("--tile-epochs", type=int, default=10, help="epochs of the multi-tile leg")
    ap.add_argument("--tile-sizes", default="128,192,256",
                    help="low-res tile sides of the multi-tile leg, cycled over the 48 tiles")
    ap.add_argument("--tiles-only", action="store_true",
                    help="only the grouped multi-tile solve, fixed streams (for profiling); prints its timing")
    ap.add_argument("--tile-streams", type=int, default=2, help="stream count of --tiles-only")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy (reference for rocprof FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch

        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read": src.numel() * 4, "bytes_written": dst.numel() * 4}),
              flush=True)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build runs on one GPU: the solve is single-device (the reference "
                                   "is single-GPU too); --gpus above 1 is not supported"}), flush=True)
        sys.exit(2)

    if args.tiles_only:
        print(json.dumps({"tiles_only": run_tiles_grouped(args)}), flush=True)
        return

    import ctypes

    import numpy as np
    import torch

    import fpm_amd
    from fpm_amd import synth

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    t_gen = time.perf_counter()
    ds = synth.make_device(args.n, args.pratio, device="cuda:0")
    print(f"[bench] synthetic data ready in {time.perf_counter() - t_gen:.1f} s "
          f"(n={ds['n']} N={ds['N']} L={ds['L']})", file=sys.stderr, flush=True)
    n, N, L = ds["n"], ds["N"], ds["L"]
    if args.y == "u16":
        Yd, y_u16 = ds["raw_dev"], True
    else:
        Yd = torch.empty(ds["raw_dev"].shape, dtype=torch.float32, device=dev)
        fpm_amd._check(fpm_amd.lib().fpm_prepare_amplitude(ds["raw_dev"].data_ptr(), ds["raw_dev"].numel(),
                                                           Yd.data_ptr(), 1, None, None, None))
        y_u16 = False
    Od = torch.empty_like(ds["O0_dev"])
    Pd = torch.empty_like(ds["P0_dev"])
    stream = torch.cuda.current_stream(dev)
    opt = (args.lr, 0.999, 1e-4)
    stitch = fpm_amd.STITCH_ATOMIC if args.stitch == "atomic" else fpm_amd.STITCH_OWNER

    def solve(epochs, stage_times=None, batch=None):
        batch = args.batch if batch is None else batch
        optv = (ctypes.c_double * 3)(*opt)
        o = fpm_amd._opts(stitch=stitch, device=True, stream=stream.cuda_stream, graph=args.graph,
                          engine=fpm_amd.FFT_SPLIT_F16 if args.engine == "split" else fpm_amd.FFT_F32)
        if stage_times is not None:
            o.stage_times = ctypes.pointer(stage_times)
        fpm_amd._check(fpm_amd.lib().fpm_pure_v3(
            ds["O0_dev"].data_ptr(), N, N, ds["P0_dev"].data_ptr(), n, n, Yd.data_ptr(),
            fpm_amd.Y_U16_RAW if y_u16 else fpm_amd.Y_F32_AMPLITUDE, L, ds["led_dev"].data_ptr(),
            ds["A_dev"].data_ptr(), float(args.pratio), float(epochs), float(batch), optv,
            Od.data_ptr(), Pd.data_ptr(), ctypes.byref(o)))

    if args.warmup > 0:
        solve(args.warmup)
        torch.cuda.synchronize()
    st = None if (args.no_stages or args.graph) else fpm_amd.StageTimes()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solve(args.steps, st)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    sec_per_epoch = (t1 - t0) / args.steps
    print(f"[bench] {args.steps} epochs in {t1 - t0:.3f} s", file=sys.stderr, flush=True)
    assert torch.isfinite(torch.view_as_real(Od)).all().item(), "non-finite object"

    steps_per_epoch = (L + args.batch - 1) // args.batch
    b_avg = L / steps_per_epoch
    y_bytes = 2 if y_u16 else 4
    epoch_bytes = steps_per_epoch * 104.0 * N * N + L * (136.0 + y_bytes) * n * n
    out = {
        "metric": METRIC, "value": sec_per_epoch, "unit": "s/epoch", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": sec_per_epoch * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"FPM solve, {L} uint16 images {n}x{n} -> {N}x{N} complex64 object, "
                               f"batch {args.batch}, {args.stitch} stitch, {args.steps} epochs in one call",
                   "n": n, "N": N, "L": L, "batch": args.batch, "y": args.y, "stitch": args.stitch,
                   "launch": "hipGraph per epoch" if args.graph else "eager",
                   "fft_engine": "split-fp16 MFMA (hi+lo, fp32 accumulate)" if args.engine == "split" else "fp32",
                   "solve_wall_s": t1 - t0,
                   "reference_wall_s_rtx3090_full_solve": 22.31},
        "epoch_dense_model_GB": epoch_bytes / 1e9,
        "epoch_dense_model_GBps": epoch_bytes / 1e9 / sec_per_epoch,
    }
    if st is not None:
        alg = stage_alg_bytes(n, N, b_avg, y_bytes, args.stitch)
        if int(st.count[STAGE_NAMES.index("crop")]) == 0:
            # the crop runs inside the first inverse column pass
            alg["ifft_n_cols"] += alg["crop"]
        if int(st.count[STAGE_NAMES.index("ifft_n_contig")]) == 0:
            # the contiguous passes run inside the stencil kernel
            alg["stencil"] += alg["ifft_n_contig"] + alg["fft_n_contig"]
        win = window_rows_per_step(ds["led"], n, N, args.batch)
        need = stage_necessary_bytes(n, N, b_avg, y_bytes, win)
        stages = {}
        for k, name in enumerate(STAGE_NAMES):
            cnt = int(st.count[k])
            if cnt == 0:
                continue
            ms_exec = st.ms[k] / cnt
            stages[name] = {"ms_total": round(st.ms[k], 4), "execs": cnt, "launches": int(st.launches[k]),
                            "ms_per_exec": round(ms_exec, 5),
                            "dense_model_GBps": round(alg[name] / 1e9 / (ms_exec / 1e3), 1)}
            if name in need:
                stages[name]["necessary_GBps"] = round(need[name] / 1e9 / (ms_exec / 1e3), 1)
        out["stages"] = stages
        dom = max(stages, key=lambda s: stages[s]["ms_total"])
        ms_dom = stages[dom]["ms_per_exec"]
        # achieved: the bytes the dominant kernel must move per launch (no
        # credit for absorbed stages) over its event-timed launch duration
        ach = need[dom] / 1e9 / (ms_dom / 1e3) if dom in need else None
        traffic = measured_traffic(dom, n, N)
        out["roofline"] = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1) if ach else None,
                           "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                           "frac": round(ach / HBM_PEAK_GBPS, 4) if ach else None, "traffic": traffic,
                           "necessary_bytes_per_exec": need.get(dom), "window_rows_per_step": win,
                           "ms_per_exec": ms_dom, "dense_model_bytes_per_exec": alg[dom],
                           "dense_model_GBps": stages[dom]["dense_model_GBps"]}
        if traffic:
            out["roofline"]["traffic_GBps"] = round(traffic / 1e9 / (ms_dom / 1e3), 1)
        # the measured denominator beside the spec peak: a plain device copy
        cp = copy_bandwidth_gbps()
        out["roofline"]["copy_GBps"] = round(cp, 1)
        if ach:
            out["roofline"]["frac_of_copy"] = round(ach / cp, 4)
    if args.batch_sweep:
        # the full-scale configuration's batch sizes: same call, same number of epochs, launched
        # without per-stage events (so every size, the bench line's included, is timed alike)
        sweep = {}
        for bs in (int(v) for v in args.batch_sweep.split(",")):
            solve(1, batch=bs)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            solve(args.steps, batch=bs)
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / args.steps
            spe = (L + bs - 1) // bs
            sweep[str(bs)] = {"s_per_epoch": dt, "steps_per_epoch": spe,
                              "epoch_dense_model_GBps": (spe * 104.0 * N * N + L * (136.0 + y_bytes) * n * n) / 1e9 / dt}
            print(f"[bench] batch {bs}: {dt * 1e3:.1f} ms/epoch", file=sys.stderr, flush=True)
        out["batch_sweep"] = sweep
    if args.batch_sweep and args.batch_autotune:
        # the opt-in minibatch autotuner (a batch-size study, so it runs with the batch
        # sweep): time to the batch-26 solve's final loss per candidate size
        from fpm_amd import tune

        cands = tuple(int(v) for v in args.batch_autotune.split(","))
        best, rep = tune.autotune_batch(ds["O0_dev"].data_ptr(), ds["P0_dev"].data_ptr(), Yd.data_ptr(), y_u16,
                                        ds["led_dev"].data_ptr(), ds["A_dev"].data_ptr(), (N, N), (n, n), L,
                                        args.pratio, opt, Od.data_ptr(), Pd.data_ptr(), candidates=cands,
                                        ref_batch=args.batch, epochs=args.steps, stream=stream.cuda_stream)
        out["batch_autotune"] = {"chosen": best, "target_loss": rep[args.batch]["losses"][-1],
                                 "candidates": {str(b): {"s_to_target": round(r["s"], 4), "epochs": r["epochs"],
                                                         "reached": r["reached"]} for b, r in sorted(rep.items())}}
        print(f"[bench] batch autotune: chosen {best}", file=sys.stderr, flush=True)
    if not args.no_fft_reference and (n, N) == (2048, 16384):
        out["fft_reference"] = run_fft_reference(stream)
        print(f"[bench] fft vs rocFFT: {out['fft_reference']}", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (oracle: one epoch, and the two smaller cases)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = run_cpu_baseline(ds, args)
    if not (args.no_time_to_loss and args.no_tiles):
        del ds, Yd, Od, Pd
        torch.cuda.empty_cache()
    if not args.no_time_to_loss:
        out["time_to_loss"] = run_time_to_loss(args)
    if not args.no_tiles:
        out["tiles"] = run_tiles(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
