"""The optional mex wrappers (fused-fourier-ptycho_amd/mex/) built against a
MATLAB API stub (tests/mex_stub/, declarations written here, host arrays)
and called as MATLAB would call them: the reference's argument checks and
outputs (cuFPM_v3/cuFPM_pure_v3.cu:44-228, benchmark/cuFPM/fullyfusedFPM.cu:
21-162), every gpuArray view released on every error path, MATLAB's GPU
library initialised only when a gpuArray is passed (gpuArray is CUDA-only,
so on an MI355X host every input is a host array), and -- on the GPU -- the
same bits as the C-ABI."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import fpm_amd
from helpers import margin_problem

HERE = os.path.dirname(os.path.abspath(__file__))
STUB = os.path.join(HERE, "mex_stub")
_libs = {}


def _wrapper(name):
    """ctypes handle of build/libmex_<name>.so (made by tests/mex_stub/Makefile)."""
    if name not in _libs:
        so = os.path.join(STUB, "build", f"libmex_{name}.so")
        if not os.path.exists(so):
            subprocess.run(["make", "-s", "-C", STUB], check=True)
        L = ctypes.CDLL(so)
        vp, i32, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
        L.stub_create.argtypes = [i32, i32, i32, ctypes.POINTER(sz), vp, i32]
        L.stub_create.restype = vp
        L.stub_data.argtypes = [vp]
        L.stub_data.restype = vp
        L.stub_bytes.argtypes = [vp]
        L.stub_bytes.restype = sz
        L.stub_ndims.argtypes = [vp]
        L.stub_dim.argtypes = [vp, i32]
        L.stub_dim.restype = sz
        L.stub_destroy.argtypes = [vp]
        L.stub_call.argtypes = [i32, ctypes.POINTER(vp), i32, ctypes.POINTER(vp)]
        L.stub_error_id.restype = ctypes.c_char_p
        L.stub_error_msg.restype = ctypes.c_char_p
        L.stub_class_id.argtypes = [ctypes.c_char_p]
        _libs[name] = L
    return _libs[name]


_CLS = {np.dtype(np.float64): "double", np.dtype(np.float32): "single", np.dtype(np.complex64): "single",
        np.dtype(np.int32): "int32", np.dtype(np.uint16): "uint16"}


def _mx(L, a, gpu=False):
    a = np.asfortranarray(a)
    dims = (ctypes.c_size_t * max(a.ndim, 2))(*(a.shape if a.ndim >= 2 else (a.size, 1)))
    cls = L.stub_class_id(_CLS[a.dtype].encode())
    return L.stub_create(cls, 1 if np.iscomplexobj(a) else 0, len(dims), dims, a.ctypes.data, 1 if gpu else 0)


def _call(L, nlhs, args):
    """call mexFunction; returns (outputs as mxArray pointers, error id, message)"""
    prhs = (ctypes.c_void_p * len(args))(*args)
    plhs = (ctypes.c_void_p * max(nlhs, 1))()
    rc = L.stub_call(nlhs, plhs, len(args), prhs)
    for a in args:
        L.stub_destroy(a)
    if rc:
        return None, L.stub_error_id().decode(), L.stub_error_msg().decode()
    return list(plhs)[:nlhs], None, None


def _out(L, p, shape, dtype):
    buf = (ctypes.c_char * L.stub_bytes(p)).from_address(L.stub_data(p))
    a = np.frombuffer(bytes(buf), dtype=dtype).reshape(shape, order="F").copy(order="F")
    L.stub_destroy(p)
    return a


def _v3_args(L, O, P, Y, led, A, r, epochs, batch, opt, gpu=(False,) * 5):
    return [_mx(L, O, gpu[0]), _mx(L, P, gpu[1]), _mx(L, Y, gpu[2]), _mx(L, led, gpu[3]), _mx(L, A, gpu[4]),
            _mx(L, np.array([[float(r)]])), _mx(L, np.array([[float(epochs)]])), _mx(L, np.array([[float(batch)]])),
            _mx(L, np.asarray(opt, dtype=np.float64).reshape(1, -1))]


def _small():
    O, P, Y, led = margin_problem(16, 2, 3)
    A = np.asfortranarray((np.abs(P) > 0).astype(np.float32))
    return O, P, Y, led, A


def test_wrappers_compile_against_the_api_stub():
    """Both wrappers compile with -Wall -Wextra against the stub headers and
    link to libfpm_amd.so (tests/mex_stub/Makefile)."""
    subprocess.run(["make", "-s", "-B", "-C", STUB], check=True)
    for name in ("cuFPM_pure_v3", "fullyfusedFPM"):
        assert os.path.exists(os.path.join(STUB, "build", f"libmex_{name}.so"))


@pytest.mark.parametrize("case,err", [
    ("nrhs", "fpm:nrhs"), ("nlhs", "fpm:nlhs"), ("real_object", "fpm:type"), ("double_led", "fpm:type"),
    ("short_opt", "fpm:opt"), ("mixed_gpu", "fpm:mixed"), ("gpu_type", "fpm:type")])
def test_cuFPM_pure_v3_wrapper_errors_release_every_view(case, err):
    L = _wrapper("cuFPM_pure_v3")
    O, P, Y, led, A = _small()
    views0, arrays0, init0 = L.stub_live_views(), L.stub_live_arrays(), L.stub_init_gpu_calls()
    nlhs, gpu, opt = 2, (False,) * 5, (1e-3, 0.999, 1e-4)
    if case == "real_object":
        O = np.real(O).astype(np.float32)
    elif case == "double_led":
        led = led.astype(np.float64)
    elif case == "short_opt":
        opt = (1e-3, 0.999)
    elif case == "mixed_gpu":
        gpu = (True, True, False, False, False)
    elif case == "gpu_type":
        gpu, O = (True,) * 5, np.real(O).astype(np.float32)
    elif case == "nlhs":
        nlhs = 3
    args = _v3_args(L, O, P, Y, led, A, 2, 1, 9, opt, gpu)
    if case == "nrhs":
        L.stub_destroy(args.pop())
    out, eid, msg = _call(L, nlhs, args)
    assert out is None and eid == err, (eid, msg)
    assert L.stub_live_views() == views0, "a gpuArray view leaked"
    assert L.stub_live_arrays() == arrays0, "an array leaked"
    assert L.stub_init_gpu_calls() - init0 == (1 if any(gpu) else 0)


def test_fullyfusedFPM_wrapper_errors_release_every_view():
    L = _wrapper("fullyfusedFPM")
    O, P, Y, led, _ = _small()
    views0, arrays0 = L.stub_live_views(), L.stub_live_arrays()
    for gpu, O_, err in (((True, False, False, False), O, "fpm:mixed"),
                         ((True, True, True, True), np.real(O).astype(np.float32), "fpm:type")):
        args = [_mx(L, O_, gpu[0]), _mx(L, P, gpu[1]), _mx(L, Y, gpu[2]), _mx(L, led, gpu[3]),
                _mx(L, np.array([[2.0]]))]
        out, eid, msg = _call(L, 2, args)
        assert eid == err, (eid, msg)
    assert L.stub_live_views() == views0 and L.stub_live_arrays() == arrays0


def test_wrapper_without_a_device_raises_the_library_error():
    """With host arrays and valid types the wrapper calls the library; on a
    machine without a gfx950 device that returns an error whose text the
    wrapper raises (fpm:solve), and both preallocated outputs are freed."""
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a device is present: the solve runs (GPU test below)")
    L = _wrapper("cuFPM_pure_v3")
    arrays0, init0 = L.stub_live_arrays(), L.stub_init_gpu_calls()
    out, eid, msg = _call(L, 2, _v3_args(L, *_small(), 2, 1, 9, (1e-3, 0.999, 1e-4)))
    print(f"[mex] {eid}: {msg}")
    assert eid == "fpm:solve" and msg
    assert L.stub_live_arrays() == arrays0
    assert L.stub_init_gpu_calls() == init0  # host arrays never initialise MATLAB's GPU library


@pytest.mark.gpu
def test_cuFPM_pure_v3_wrapper_equals_the_c_abi():
    """[O, P] = cuFPM_pure_v3(O0, P0, Y, ledIdx, A, pratio, epochs, batch,
    [lr beta eps]) through the wrapper: the same bits as fpm_pure_v3 called
    directly (ragged last batch, 2 epochs)."""
    L = _wrapper("cuFPM_pure_v3")
    O, P, Y, led, A = _small()
    opt = (1e-3, 0.999, 1e-4)
    init0 = L.stub_init_gpu_calls()
    out, eid, msg = _call(L, 2, _v3_args(L, O, P, Y, led, A, 2, 2, 4, opt))
    assert eid is None, (eid, msg)
    Ow = _out(L, out[0], O.shape, np.complex64)
    Pw = _out(L, out[1], P.shape, np.complex64)
    Oc, Pc = fpm_amd.cuFPM_pure_v3(O, P, Y, led, A, 2, 2, 4, opt)
    assert np.array_equal(Ow, Oc) and np.array_equal(Pw, Pc)
    assert L.stub_init_gpu_calls() == init0


@pytest.mark.gpu
def test_fullyfusedFPM_wrapper_equals_the_c_abi():
    L = _wrapper("fullyfusedFPM")
    O, P, Y, led, _ = _small()
    args = [_mx(L, O), _mx(L, P), _mx(L, Y), _mx(L, led), _mx(L, np.array([[2.0]]))]
    out, eid, msg = _call(L, 2, args)
    assert eid is None, (eid, msg)
    G1 = _out(L, out[0], O.shape, np.complex64)
    G2 = _out(L, out[1], P.shape, np.complex64)
    g1, g2, _ = fpm_amd.fused_gradient(O, P, Y, led, 2)
    assert np.array_equal(G1, g1) and np.array_equal(G2, g2)
