"""Minibatch size chosen by measured time to loss (opt-in).

The minibatch size is part of the method, not of the schedule: the
reference walks the LEDs in contiguous chunks of `batch`
(cuFPM_v3/cuFPM_pure_v3.cu:129-141), so another size is another trajectory.
The library therefore never changes it by itself (fpm_tiles_solve tunes only
the overlap of its launch sequences).  autotune_batch is for callers who
accept any size that reaches the same data loss: the reference size runs its
epochs and its final epoch's loss becomes the target; every other candidate
runs until the first epoch whose loss is at or below that target
(fpm_opts.stop_loss), at most `max_epochs` epochs, each timed with device
synchronisation; the candidate that reaches the target soonest is chosen.
"""
from __future__ import annotations

import time
from typing import Dict, Sequence, Tuple

from . import STITCH_OWNER, solve_device


def autotune_batch(O0_ptr: int, P0_ptr: int, Y_ptr: int, y_u16: bool, led_ptr: int, A_ptr: int,
                   N: Tuple[int, int], n: Tuple[int, int], L: int, pratio: float, opt: Sequence[float],
                   O_out_ptr: int, P_out_ptr: int, *, candidates: Sequence[int] = (16, 26, 32),
                   ref_batch: int = 26, epochs: int = 20, max_epochs: int = 0, stitch: int = STITCH_OWNER,
                   engine: int = 0, stream=None) -> Tuple[int, Dict[int, dict]]:
    """(chosen batch, report).  Arguments as solve_device (device pointers,
    column-major buffers in HBM); O_out / P_out receive each trial in turn
    (the last trial's result is left there).  report[b] holds the seconds to
    the target, the epochs run, the per-epoch losses and whether the target
    was reached.  max_epochs (0: epochs) caps the other candidates' trials."""
    import torch

    cap = int(max_epochs) if max_epochs > 0 else int(epochs)

    def trial(b: int, ne: int, stop: float):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        losses = solve_device(O0_ptr, P0_ptr, Y_ptr, y_u16, led_ptr, A_ptr, N, n, L, pratio, ne, b, opt, O_out_ptr,
                              P_out_ptr, stitch=stitch, engine=engine, stream=stream, loss=True, stop_loss=stop)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, [float(v) for v in losses]

    trial(int(ref_batch), 1, 0.0)  # first-call costs (code objects, tables) out of the timings
    t_ref, l_ref = trial(int(ref_batch), int(epochs), 0.0)
    target = l_ref[-1]
    report = {int(ref_batch): {"s": t_ref, "epochs": len(l_ref), "losses": l_ref, "reached": True}}
    for b in candidates:
        b = int(b)
        if b in report:
            continue
        t, l = trial(b, cap, target)
        report[b] = {"s": t, "epochs": len(l), "losses": l, "reached": bool(l) and l[-1] <= target}
    best = min((b for b in report if report[b]["reached"]), key=lambda b: report[b]["s"])
    return best, report
