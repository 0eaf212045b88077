"""fpm_fused_gradient_noalloc (the gradient entry whose working buffers live in
a static device array instead of allocations) against fpm_fused_gradient
with the same options (device pointers, FP32 transforms, owner stitch, float
amplitudes): the same bits on two calls in a row, the second smaller (the
array keeps the first call's contents), the loss slot cleared by the entry,
and a call that would need more than the array holds rejected."""
import ctypes

import numpy as np
import pytest

import fpm_amd
from helpers import margin_problem

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _dev(a):
    """host array (MATLAB orientation) -> flat device buffer in column-major order"""
    return torch.from_numpy(np.ravel(np.asarray(a), order="F").copy()).cuda()


def _host(t, shape):
    return np.asfortranarray(t.cpu().numpy().reshape(shape, order="F"))


def _noalloc(O, P, Y, led, r, loss_init=123.0):
    L = fpm_amd.lib()
    Od, Pd, Yd, ld = _dev(O), _dev(P), _dev(Y), _dev(led)
    GO, GP = torch.empty_like(Od), torch.empty_like(Pd)
    loss = torch.full((1,), loss_init, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    torch.cuda.synchronize()
    rc = L.fpm_fused_gradient_noalloc(Od.data_ptr(), O.shape[0], O.shape[1], Pd.data_ptr(), P.shape[0], P.shape[1],
                                      Yd.data_ptr(), Y.shape[2], ld.data_ptr(), float(r), GO.data_ptr(),
                                      GP.data_ptr(), loss.data_ptr(), ctypes.c_void_p(stream))
    assert rc == 0, L.fpm_last_error().decode()
    torch.cuda.synchronize()
    return _host(GO, O.shape), _host(GP, P.shape), float(loss.item())


def _reference(O, P, Y, led, r):
    Od, Pd, Yd, ld = _dev(O), _dev(P), _dev(Y), _dev(led)
    GO, GP = torch.empty_like(Od), torch.empty_like(Pd)
    torch.cuda.synchronize()
    loss = fpm_amd.gradient_device(Od.data_ptr(), Pd.data_ptr(), Yd.data_ptr(), False, ld.data_ptr(), O.shape,
                                   P.shape, Y.shape[2], r, GO.data_ptr(), GP.data_ptr())
    torch.cuda.synchronize()
    return _host(GO, O.shape), _host(GP, P.shape), loss


def test_noalloc_gradient_equals_fused_gradient_on_two_calls():
    """256 -> 1024 with 25 images (44 MB of the 64 MiB array), then 64 ->
    256 with 9 images: each equal to fpm_fused_gradient bit for bit."""
    for n, r, grid in ((256, 4, 5), (64, 4, 3)):
        print(f"[noalloc] n={n} r={r} L={grid * grid}", flush=True)
        O, P, Y, led = margin_problem(n, r, grid)
        a = _noalloc(O, P, Y, led, r)
        b = _reference(O, P, Y, led, r)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (n, r)
        assert abs(a[2] - b[2]) <= 1e-12 * b[2], (a[2], b[2])


def test_noalloc_rejects_a_call_beyond_its_scratch():
    """512 -> 2048 needs 2 x 32 MiB of spectra alone: FPM_E_INVALID, nothing launched."""
    L = fpm_amd.lib()
    buf = torch.zeros(16, dtype=torch.complex64, device="cuda")
    p = buf.data_ptr()
    rc = L.fpm_fused_gradient_noalloc(p, 2048, 2048, p, 512, 512, p, 26, p, 4.0, p, p, None, None)
    assert rc == -1 and "scratch" in L.fpm_last_error().decode()
