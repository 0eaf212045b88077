"""The opt-in minibatch autotuner (fpm_amd.tune.autotune_batch): the
reference size runs its epochs, every other candidate runs until its first
epoch at or below the reference's final loss, and the size that gets there
soonest is chosen."""
import numpy as np
import pytest

import fpm_amd
from fpm_amd import synth, tune

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_autotune_batch_picks_the_fastest_size_to_the_reference_loss():
    """64^2 x 49 uint16 images -> 256^2: reference batch 13 for 8 epochs,
    candidates 7 and 25 capped at 16 epochs.  Every trial that reaches the
    target stops at its first epoch at or below it, the chosen size has the
    least time among those, and its trial repeated gives the same per-epoch
    losses bit for bit (owner stitch: deterministic)."""
    ds = synth.make_device(64, 4, led_grid=7)
    n, N, L = ds["n"], ds["N"], ds["L"]
    O, P = torch.empty_like(ds["O0_dev"]), torch.empty_like(ds["P0_dev"])
    opt = (5e-4, 0.999, 1e-4)
    args = (ds["O0_dev"].data_ptr(), ds["P0_dev"].data_ptr(), ds["raw_dev"].data_ptr(), True, ds["led_dev"].data_ptr(),
            ds["A_dev"].data_ptr(), (N, N), (n, n), L, 4)
    best, rep = tune.autotune_batch(*args, opt, O.data_ptr(), P.data_ptr(), candidates=(7, 13, 25), ref_batch=13,
                                    epochs=8, max_epochs=16)
    print(f"[tune] chosen {best}: " + ", ".join(f"{b}: {r['epochs']} ep {r['s'] * 1e3:.1f} ms reached {r['reached']}"
                                               for b, r in sorted(rep.items())), flush=True)
    assert sorted(rep) == [7, 13, 25]
    target = rep[13]["losses"][-1]
    assert rep[13]["epochs"] == 8 and rep[13]["reached"]
    for b, r in rep.items():
        if b == 13:
            continue
        if r["reached"]:
            assert r["losses"][-1] <= target and r["epochs"] <= 16
            assert all(v > target for v in r["losses"][:-1]), b
        else:
            assert r["epochs"] == 16 and min(r["losses"]) > target, b
    assert rep[best]["reached"]
    assert rep[best]["s"] == min(r["s"] for r in rep.values() if r["reached"])
    ne, stop = (8, 0.0) if best == 13 else (16, target)
    again = fpm_amd.solve_device(*args, ne, best, opt, O.data_ptr(), P.data_ptr(), loss=True, stop_loss=stop)
    assert np.array_equal(again, np.asarray(rep[best]["losses"]))
