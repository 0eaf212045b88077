"""Field-of-view tiles reconstructed jointly (fpm_tiles_solve).

No counterpart in the reference, whose mex reconstructs one field of view
seen from its centre.  Here a large field is split into overlapping tiles of
several sizes; every tile has its own pupil (aberrations vary across the
field) and its own LED windows (each LED is seen under a different angle from
each tile centre: fpm_led_offsets_at), and all tiles run through one hipGraph
per epoch on a few concurrent streams.

make_tile_set generates such a field synthetically on the GPU: one object
over the whole field (the generator of fpm_amd.synth: band-limited amplitude
in [0.3, 1], phase from Gaussians), tiles on a grid with sizes cycling
through `sizes` (powers of two), per-tile pupils A exp(i Z) with Zernike
defocus, astigmatism (0 and 45 degrees) and coma (x and y) coefficients drawn
from N(0, sigma) rad, images through the library's image model, Poisson
noise, uint16.
"""
from __future__ import annotations

import ctypes
from typing import List, Sequence

import numpy as np

from . import _check, lib

_bound = False


class Tile(ctypes.Structure):
    _fields_ = [("O0", ctypes.c_void_p), ("O_out", ctypes.c_void_p), ("N", ctypes.c_int64),
                ("P0", ctypes.c_void_p), ("P_out", ctypes.c_void_p), ("n", ctypes.c_int64),
                ("Y", ctypes.c_void_p), ("y_dtype", ctypes.c_int32), ("led_idx", ctypes.c_void_p),
                ("A", ctypes.c_void_p), ("loss_per_epoch", ctypes.POINTER(ctypes.c_double))]


# fpm_opts.tile_launch
TILES_GROUPED = 0   # tiles of one size and image type share launches
TILES_PER_TILE = 1  # one launch sequence per tile


class TilesReport(ctypes.Structure):
    _fields_ = [("streams", ctypes.c_int32), ("tuned", ctypes.c_int32), ("cand", ctypes.c_int32 * 4),
                ("cand_ms", ctypes.c_double * 4), ("sequences", ctypes.c_int32), ("grouped_tiles", ctypes.c_int32)]


def _tiles_lib():
    global _bound
    from . import _Opts

    L = lib()
    if not _bound:
        vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        L.fpm_tiles_solve.argtypes = [ctypes.POINTER(Tile), i32, i64, dbl, dbl, dbl, ctypes.POINTER(dbl), i32,
                                      ctypes.POINTER(_Opts), ctypes.POINTER(TilesReport)]
        L.fpm_tiles_solve.restype = ctypes.c_int
        L.fpm_led_offsets_at.argtypes = [dbl, dbl, dbl, dbl, dbl, dbl, dbl, i32, i64, i32, i32, vp, vp, vp]
        L.fpm_led_offsets_at.restype = ctypes.c_int
        _bound = True
    return L


def led_offsets_at(objdx_um: float, objdy_um: float, n: int, pratio: int, *, wavelength_um=0.532, na=0.1,
                   pixel_um=1.625, led_pitch_um=4000.0, led_height_um=90000.0, led_grid=19, spiral=True):
    """(led_idx int32 2 x L Fortran, grid_index, na_illum) of the field area
    at (objdx, objdy) um (init_environment_rgb.m:26-29,83-84)."""
    L = led_grid * led_grid
    led = np.zeros((2, L), dtype=np.int32, order="F")
    gi = np.zeros(L, dtype=np.int32)
    nai = np.zeros(L, dtype=np.float64)
    _check(_tiles_lib().fpm_led_offsets_at(float(objdx_um), float(objdy_um), wavelength_um, na, pixel_um, led_pitch_um,
                                           led_height_um, led_grid, n, pratio, 1 if spiral else 0, led.ctypes.data,
                                           gi.ctypes.data, nai.ctypes.data))
    return led, gi, nai


def tiles_solve(tiles: Sequence[dict], pratio: int, epochs: int, batch: int, opt, *, streams: int = 0,
                stitch=None, engine: int = 0, loss: bool = False, stream=None, tile_launch: int = TILES_GROUPED):
    """Jointly solve device-resident tiles.  Each tile dict holds CUDA
    tensors in the solve's device layout: O0, O (output), P0, P (output),
    Y (uint16 raw or float32 amplitudes), led (int32 L x 2 rows = 2 x L
    column-major), A.  tile_launch: TILES_GROUPED (tiles of one size share
    launches) or TILES_PER_TILE.  Returns (report dict, per-tile loss arrays
    or None)."""
    import torch

    from . import STITCH_OWNER, Y_F32_AMPLITUDE, Y_U16_RAW, _opts

    arr = (Tile * len(tiles))()
    losses = []
    Ls = {int(t["Y"].shape[0]) for t in tiles}
    if len(Ls) != 1:
        raise ValueError("every tile must hold the same number of images")
    L = Ls.pop()
    for k, t in enumerate(tiles):
        lb = np.zeros(max(int(epochs), 1), dtype=np.float64) if loss else None
        losses.append(lb)
        arr[k] = Tile(t["O0"].data_ptr(), t["O"].data_ptr(), int(t["O0"].shape[0]), t["P0"].data_ptr(),
                      t["P"].data_ptr(), int(t["P0"].shape[0]), t["Y"].data_ptr(),
                      Y_U16_RAW if t["Y"].dtype == torch.uint16 else Y_F32_AMPLITUDE, t["led"].data_ptr(),
                      t["A"].data_ptr(), lb.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if loss else None)
    o = _opts(stitch=STITCH_OWNER if stitch is None else stitch, device=True, stream=stream, engine=engine,
              tile_launch=tile_launch)
    rep = TilesReport()
    optv = (ctypes.c_double * 3)(*[float(v) for v in opt[:3]])
    torch.cuda.synchronize()
    _check(_tiles_lib().fpm_tiles_solve(arr, len(tiles), L, float(pratio), float(epochs), float(batch), optv,
                                        int(streams), ctypes.byref(o), ctypes.byref(rep)))
    report = {"streams": rep.streams, "tuned": rep.tuned,
              "candidates": {int(rep.cand[k]): float(rep.cand_ms[k]) for k in range(rep.tuned)},
              "sequences": rep.sequences, "grouped_tiles": rep.grouped_tiles}
    return report, ([lb[:int(epochs)] for lb in losses] if loss else None)


# --------------------------------------------------------- generator
def make_tile_set(field_n: int, sizes: Sequence[int], grid: int, pratio: int, *, seed: int = 7, sigma: float = 0.4,
                  led_grid: int = 19, count: int = 0, device: str = "cuda") -> List[dict]:
    """Tiles of a synthetic field of field_n x field_n low-res pixels
    (field_n a power of two: the field object is made with the library's
    transforms): a
    grid x grid lattice of tile centres (the first `count` of them, all if
    0), sizes cycling through `sizes`; returns tile dicts for tiles_solve
    plus the true pupil ("P_true") and window ("window": low-res (row0, col0))."""
    import torch
    import torch.nn.functional as F

    from . import _check, lib, synth

    dev = torch.device(device)
    stream = torch.cuda.current_stream(dev).cuda_stream
    NF = field_n * pratio
    rng = np.random.Generator(np.random.Philox(seed))
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    # field object (the generator of synth.make_device over the whole field)
    noise = torch.rand((NF, NF), generator=gen, device=dev, dtype=torch.float32).to(torch.complex64)
    spec = torch.empty_like(noise)
    _check(lib().fpm_fft2(noise.data_ptr(), spec.data_ptr(), NF, NF, 1, -1, 1, stream))
    f = torch.fft.fftfreq(NF, d=synth.PIXEL_UM / pratio, device=dev, dtype=torch.float64) * 2 * np.pi
    _, _, nai = synth.led_offsets(64, pratio, led_grid=led_grid)
    cut = (synth.NA + float(nai.max())) * 2 * np.pi / synth.WAVELENGTH_UM
    spec *= ((f[:, None] ** 2 + f[None, :] ** 2) <= cut * cut).to(torch.float32)
    _check(lib().fpm_fft2(spec.data_ptr(), noise.data_ptr(), NF, NF, 1, 1, 1, stream))
    amp = noise.real.clone()
    del noise, spec
    amp = 0.3 + 0.7 * (amp - amp.min()) / (amp.max() - amp.min())
    phase = torch.zeros((NF, NF), device=dev, dtype=torch.float32)
    ax = torch.arange(NF, device=dev, dtype=torch.float32)
    for _ in range(20 * max(1, (field_n // 256) ** 2)):
        cy, cx = rng.random(2) * NF
        sig = NF / 64 + rng.random() * (NF / 16 - NF / 64)
        h = (rng.random() * 2 - 1) * np.pi / 2
        phase.add_(h * torch.outer(torch.exp(-((ax - float(cx)) ** 2) / (2 * sig * sig)),
                                   torch.exp(-((ax - float(cy)) ** 2) / (2 * sig * sig))))
    O_field = torch.polar(amp, phase)
    del amp, phase
    tiles = []
    centres = [((gy + 0.5) * field_n / grid, (gx + 0.5) * field_n / grid) for gy in range(grid) for gx in range(grid)]
    if count:
        centres = centres[:count]
    for k, (cy, cx) in enumerate(centres):
        n = int(sizes[k % len(sizes)])
        N = n * pratio
        r0 = int(min(max(round(cy - n / 2), 0), field_n - n))
        c0 = int(min(max(round(cx - n / 2), 0), field_n - n))
        # field area of the tile: its centre relative to the field centre in
        # um, entering the LED offsets with the sign of init_environment_rgb.m:83-84
        objdx = -(c0 + n / 2 - field_n / 2) * synth.PIXEL_UM
        objdy = -(r0 + n / 2 - field_n / 2) * synth.PIXEL_UM
        led, _, _ = led_offsets_at(objdx, objdy, n, pratio, led_grid=led_grid)
        A, _ = synth.pupil(n)
        P_true = synth.zernike_pupil(n, rng.normal(0.0, sigma, 5), A)
        # device layout: row-major buffer of the column-major array (t[j, i] = M(i, j))
        Ot = O_field[c0 * pratio:c0 * pratio + N, r0 * pratio:r0 * pratio + N].contiguous()
        Pt = torch.from_numpy(np.ascontiguousarray(P_true.T)).to(dev)
        ledd = torch.from_numpy(np.ascontiguousarray(led.T)).to(dev)
        L = led.shape[1]
        I = torch.empty((L, n, n), dtype=torch.float32, device=dev)
        _check(lib().fpm_forward_intensity(Ot.data_ptr(), N, N, Pt.data_ptr(), n, n, L, ledd.data_ptr(),
                                           float(pratio), 1, I.data_ptr(), 1, stream))
        I *= synth.PEAK_COUNTS / float(I[0].max())
        I = torch.poisson(I, generator=gen) + synth.OFFSET_COUNTS
        counts = I.round_().clamp_(0, 65535).to(torch.int32)
        mn, mx = int(counts.min()), int(counts.max())
        centre = torch.sqrt((counts[0].to(torch.float32) - mn) / float(mx - mn))
        O0 = F.interpolate(centre[None, None], scale_factor=pratio, mode="bilinear",
                           align_corners=False)[0, 0].to(torch.complex64).contiguous()
        P0 = torch.from_numpy(np.ascontiguousarray(A.T.astype(np.complex64))).to(dev)
        tiles.append({"O0": O0, "O": torch.empty_like(O0), "P0": P0, "P": torch.empty_like(P0),
                      "Y": counts.to(torch.uint16), "led": ledd, "A": torch.from_numpy(np.ascontiguousarray(A.T)).to(dev),
                      "P_true": P_true, "window": (r0, c0), "n": n})
    torch.cuda.synchronize(dev)
    return tiles
