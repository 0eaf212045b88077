"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and rejects bad arguments before touching a device."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import fpm_amd
from helpers import small_leds_oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fpm_amd.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(fpm_\w+)\s*\(", text, re.M)))


def test_header_declares_the_python_symbol_list():
    assert _declared() == sorted(fpm_amd.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    lib = fpm_amd.lib()
    for s in _declared():
        assert hasattr(lib, s), s
    out = subprocess.run(["nm", "-D", "--defined-only", fpm_amd.library_path], capture_output=True, text=True).stdout
    for s in _declared():
        assert re.search(rf"\bT {s}\b", out), s


def test_abi_version_and_opts_layout():
    assert fpm_amd.lib().fpm_abi_version() == 3
    o = fpm_amd._Opts()
    fpm_amd.lib().fpm_opts_init(ctypes.byref(o))
    assert o.struct_size == ctypes.sizeof(fpm_amd._Opts)
    assert abs(o.pupil_clamp - 0.2) < 1e-7 and o.stitch == fpm_amd.STITCH_OWNER
    assert o.tile_launch == 0


def test_ctypes_mirrors_match_the_header_layout(tmp_path):
    """Every field offset and the size of fpm_opts, fpm_tile and
    fpm_tiles_report as gcc lays out the header equal the ctypes mirrors'."""
    from fpm_amd import tiles as ft
    structs = {"fpm_opts": fpm_amd._Opts, "fpm_tile": ft.Tile, "fpm_tiles_report": ft.TilesReport}
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def _call_solve(nr=16, Nr=32, L=4, batch=2, led=None, epochs=1):
    O = np.zeros((Nr, Nr), np.complex64, order="F")
    P = np.zeros((nr, nr), np.complex64, order="F")
    Y = np.ones((nr, nr, L), np.float32, order="F")
    A = np.ones((nr, nr), np.float32, order="F")
    if led is None:
        led = np.ones((2, L), np.int32, order="F")
    return fpm_amd.cuFPM_pure_v3(O, P, Y, led, A, 2, epochs, batch, (1e-3, 0.999, 1e-4))


def test_rejects_sides_outside_the_domain():
    """Every side even and in [16, 16384] (the reference takes any even side):
    odd sides, sides below 16 and above 16384 are FPM_E_UNSUPPORTED before any
    device work."""
    for nr, Nr in ((15, 30), (16, 33), (8, 16), (16, 16386), (16, 32768)):
        with pytest.raises(fpm_amd.FPMError) as e:
            _call_solve(nr=nr, Nr=Nr)
        assert e.value.code == -2, (nr, Nr)


@pytest.mark.parametrize("nr,Nr", [(24, 48), (40, 80), (20, 60), (16, 80), (134, 268), (1536, 12288)])
def test_accepts_any_even_side_up_to_the_device(nr, Nr):
    """2^a 3^b, factors 5 and 7, a prime factor 67 and 12288-point sides pass
    the size checks: without a GPU the call fails only at the device check
    (FPM_E_NODEVICE), not as unsupported."""
    try:
        _call_solve(nr=nr, Nr=Nr)
    except fpm_amd.FPMError as e:
        assert e.code == -5, e


def test_generic_geometry_rejects_split_engine_and_atomic_stitch():
    """On the generic path the split-FP16 engine and the atomic stitch are not
    available: FPM_E_UNSUPPORTED with the reason, before any device work."""
    O = np.zeros((80, 80), np.complex64, order="F")
    P = np.zeros((40, 40), np.complex64, order="F")
    Y = np.ones((40, 40, 2), np.float32, order="F")
    led = np.ones((2, 2), np.int32, order="F")
    for kw, word in (({"engine": fpm_amd.FFT_SPLIT_F16}, "split"), ({"stitch": fpm_amd.STITCH_ATOMIC}, "atomic")):
        with pytest.raises(fpm_amd.FPMError) as e:
            fpm_amd.fused_gradient(O, P, Y, led, 2, **kw)
        assert e.value.code == -2 and word in str(e.value), kw
        with pytest.raises(fpm_amd.FPMError) as e:
            fpm_amd.cuFPM_pure_v3(O, P, Y, led, np.ones((40, 40), np.float32, order="F"), 2, 1, 2,
                                  (1e-3, 0.999, 1e-4), **kw)
        assert e.value.code == -2 and word in str(e.value), kw
    with pytest.raises(fpm_amd.FPMError) as e:
        fpm_amd.fft2(O, engine=fpm_amd.FFT_SPLIT_F16)
    assert e.value.code == -2 and "split" in str(e.value)


def test_rejects_bad_batch():
    with pytest.raises(fpm_amd.FPMError) as e:
        _call_solve(batch=0)
    assert e.value.code == -1 and "batch" in str(e.value)


def test_rejects_pupil_larger_than_object():
    with pytest.raises(fpm_amd.FPMError) as e:
        _call_solve(nr=64, Nr=32)
    assert e.value.code == -1


def test_rejects_negative_epochs():
    with pytest.raises(fpm_amd.FPMError) as e:
        _call_solve(epochs=-1)
    assert e.value.code == -1


def test_led_offsets_match_oracle_restatement():
    for n, r, grid in [(128, 4, 19), (512, 8, 19), (2048, 8, 19), (64, 2, 5)]:
        a, _, _ = fpm_amd.led_offsets(n, r, led_grid=grid)
        assert np.array_equal(a, small_leds_oracle(n, r, grid)), (n, r, grid)
    a, _, _ = fpm_amd.led_offsets(128, 4, spiral=False)
    b = small_leds_oracle(128, 4, 19, spiral=False)
    assert np.array_equal(a, b)


def test_led_offsets_at_field_areas_match_oracle_restatement():
    """fpm_led_offsets_at equals the numpy restatement of
    init_environment_rgb.m:83-84 with objdx, objdy at several field areas,
    and equals fpm_led_offsets at (0, 0)."""
    import oracle
    from fpm_amd import tiles as ft
    for n, r, grid, dx, dy in [(128, 4, 19, 0.0, 0.0), (128, 4, 19, -600.5, 250.0), (64, 4, 7, 1300.0, -975.0),
                               (256, 8, 19, -2000.0, -1500.25)]:
        a, _, nai = ft.led_offsets_at(dx, dy, n, r, led_grid=grid)
        b, _, _ = oracle.led_offsets_matlab(n, r, led_grid=grid, objdx=dx, objdy=dy)
        assert np.array_equal(a, b), (n, r, grid, dx, dy)
        if dx == 0.0 and dy == 0.0:
            c, _, _ = fpm_amd.led_offsets(n, r, led_grid=grid)
            assert np.array_equal(a, c)


def test_led_offsets_rejects_even_grid():
    with pytest.raises(fpm_amd.FPMError):
        fpm_amd.led_offsets(128, 4, led_grid=4)


def test_opts_new_fields_default_and_bad_values():
    """graph / fft_engine / lr_halve_from / stop_loss default to the
    reference's behaviour; an unknown engine, a negative schedule and a
    negative or NaN stop loss are rejected before any device work."""
    o = fpm_amd._Opts()
    fpm_amd.lib().fpm_opts_init(ctypes.byref(o))
    assert o.graph == 0 and o.fft_engine == fpm_amd.FFT_F32 and o.lr_halve_from == 0
    Oz = np.zeros((32, 32), np.complex64, order="F")
    Pz = np.zeros((16, 16), np.complex64, order="F")
    Y = np.ones((16, 16, 4), np.float32, order="F")
    led = np.ones((2, 4), np.int32, order="F")
    A = np.ones((16, 16), np.float32, order="F")
    assert o.stop_loss == 0.0 and not o.epochs_run
    for field, val, word in (("fft_engine", 7, "fft_engine"), ("lr_halve_from", -1, "lr_halve_from"),
                             ("stop_loss", -1.0, "stop_loss"), ("stop_loss", float("nan"), "stop_loss")):
        o = fpm_amd._opts()
        setattr(o, field, val)
        optv = (ctypes.c_double * 3)(1e-3, 0.999, 1e-4)
        Oo, Po = np.empty_like(Oz), np.empty_like(Pz)
        rc = fpm_amd.lib().fpm_pure_v3(Oz.ctypes.data, 32, 32, Pz.ctypes.data, 16, 16, Y.ctypes.data, 0, 4,
                                       led.ctypes.data, A.ctypes.data, 2.0, 1.0, 2.0, optv, Oo.ctypes.data,
                                       Po.ctypes.data, ctypes.byref(o))
        assert rc == -1 and word in fpm_amd.lib().fpm_last_error().decode()


def test_tiles_solve_rejects_bad_arguments():
    from fpm_amd import tiles as ft
    L = ft._tiles_lib()
    optv = (ctypes.c_double * 3)(1e-3, 0.999, 1e-4)
    arr = (ft.Tile * 1)()
    # null arrays of a tile, no tiles, bad stream count: FPM_E_INVALID before any device work
    assert L.fpm_tiles_solve(arr, 1, 4, 2.0, 1.0, 2.0, optv, 0, None, None) == -1
    assert "null" in L.fpm_last_error().decode()
    assert L.fpm_tiles_solve(arr, 0, 4, 2.0, 1.0, 2.0, optv, 0, None, None) == -1
    assert L.fpm_tiles_solve(arr, 1, 4, 2.0, 1.0, 2.0, optv, 65, None, None) == -1
    # an odd tile side (outside the reference's domain of even sides)
    buf = np.zeros(4096, np.complex64)
    p = buf.ctypes.data
    arr[0] = ft.Tile(p, p, 42, p, p, 21, p, 0, p, p, None)
    assert L.fpm_tiles_solve(arr, 1, 4, 2.0, 1.0, 2.0, optv, 1, None, None) == -2
    # stop_loss is an option of the single solve only
    o = fpm_amd._opts(stop_loss=1.0)
    assert L.fpm_tiles_solve(arr, 1, 4, 2.0, 1.0, 2.0, optv, 1, ctypes.byref(o), None) == -1
    assert "stop_loss" in L.fpm_last_error().decode()
    # unknown tile launch mode
    arr[0] = ft.Tile(p, p, 32, p, p, 16, p, 0, p, p, None)
    o = fpm_amd._opts(tile_launch=2)
    assert L.fpm_tiles_solve(arr, 1, 4, 2.0, 1.0, 2.0, optv, 1, ctypes.byref(o), None) == -1
    assert "tile_launch" in L.fpm_last_error().decode()


def test_optimizer_entries_reject_bad_arguments():
    """fpm_optimizer_step / fpm_pupil_constraint: unknown kind, null arrays,
    missing moments, Adam without an iteration, negative counts:
    FPM_E_INVALID before any device work."""
    from fpm_amd import optimizers as fo
    L = fo._olib()
    buf = np.zeros(64, np.complex64)
    p = buf.ctypes.data
    assert L.fpm_optimizer_step(7, p, p, p, p, 4, 1e-3, 0.9, 0.999, 1, None) == -1
    assert "kind" in L.fpm_last_error().decode()
    assert L.fpm_optimizer_step(fo.OPT_SGD, None, p, None, None, 4, 1e-3, 0.9, 0.999, 1, None) == -1
    assert L.fpm_optimizer_step(fo.OPT_ADAM, p, p, None, None, 4, 1e-3, 0.9, 0.999, 1, None) == -1
    assert L.fpm_optimizer_step(fo.OPT_ADAM, p, p, p, p, 4, 1e-3, 0.9, 0.999, 0, None) == -1
    assert L.fpm_optimizer_step(fo.OPT_YOGI, p, p, p, p, -1, 1e-3, 0.9, 0.999, 1, None) == -1
    assert L.fpm_pupil_constraint(None, p, 4, 0.1, None) == -1
    assert L.fpm_pupil_constraint(p, p, -2, 0.1, None) == -1
    with pytest.raises(NotImplementedError):
        fo.optimizer_RMSprop(0, 0, 0.9, 0.5, False, 1e-3)
    with pytest.raises(ValueError):
        fo.optimizer_adam(1, 0, 0.9, 0.999, 1e-3)


def test_cnn_create_rejects_bad_descriptors():
    from fpm_amd import warmstart as ws
    L = ws._cnn_lib()
    h = ctypes.c_void_p()
    for desc in (ws.CnnDesc(0, 32, 26, 64, 4, 1),     # levels
                 ws.CnnDesc(4, 48, 26, 64, 4, 1),     # base not a multiple of 32
                 ws.CnnDesc(5, 32, 26, 64, 4, 1),     # 32 * 2^4 channels > 256
                 ws.CnnDesc(4, 32, 26, 12, 4, 1),     # n not divisible by 2^(levels-1)
                 ws.CnnDesc(4, 32, 26, 64, 2, 1),     # 2 up^2 not a multiple of 32
                 ws.CnnDesc(4, 32, 26, 64, 4, 0),     # max_batch
                 ws.CnnDesc(4, 32, 26, 64, 4, 8, -1, 0.1),   # pupil terms < 0
                 ws.CnnDesc(4, 32, 26, 64, 4, 8, 17, 0.1),   # pupil terms > 16
                 ws.CnnDesc(4, 32, 26, 64, 4, 65, 5, 0.1),   # batch > 64 with a pupil head
                 ws.CnnDesc(4, 32, 26, 64, 4, 8, 5, -1.0),   # negative pupil weight
                 ws.CnnDesc(4, 32, 26, 64, 4, 8, 5, float("nan"))):
        assert L.fpm_cnn_create(ctypes.byref(desc), 0, ctypes.byref(h)) == -1
        assert not h.value
    assert L.fpm_cnn_num_params(None) == -1


def test_parallax_shifts_rejects_bad_arguments():
    """fpm_parallax_shifts: null arrays, planes smaller than 3 x 3, a
    reference or plane index outside the sample's planes, more than 64
    planes, a batch of 0: FPM_E_INVALID before any device work."""
    from fpm_amd import warmstart as ws
    L = ws._cnn_lib()
    buf = np.zeros(64, np.float32)
    p = buf.ctypes.data
    which = np.array([1, 2], np.int32)
    many = np.arange(65, dtype=np.int32)
    w = which.ctypes.data
    cases = [(None, 1, 3, 8, 8, 0, w, 2, 4, p), (p, 1, 3, 8, 8, 0, None, 2, 4, p), (p, 1, 3, 8, 8, 0, w, 2, 4, None),
             (p, 1, 3, 2, 8, 0, w, 2, 4, p), (p, 1, 3, 8, 8, 3, w, 2, 4, p), (p, 1, 2, 8, 8, 0, w, 2, 4, p),
             (p, 1, 70, 8, 8, 0, many.ctypes.data, 65, 4, p), (p, 0, 3, 8, 8, 0, w, 2, 4, p),
             (p, 1, 3, 8, 8, 0, w, 2, -1, p)]
    for c in cases:
        assert L.fpm_parallax_shifts(*c, None) == -1, c
        assert "fpm_parallax_shifts" in fpm_amd.lib().fpm_last_error().decode()


def test_parallax_pupil_fit_recovers_a_linear_map():
    """ParallaxPupil's least-squares map (FP64, with a bias column) recovers
    an exact linear relation between shift features and coefficients, and
    predict applies it; the features are stubbed (no device needed)."""
    torch = pytest.importorskip("torch")
    from fpm_amd import warmstart as ws

    class Stub(ws.ParallaxPupil):
        def features(self, feats):
            return feats.reshape(feats.shape[0], -1)

    rng = np.random.default_rng(5)
    Wt = rng.standard_normal((6, 5))
    bias = rng.standard_normal(5)
    X = rng.standard_normal((300, 6))
    Z = X @ Wt + bias
    m = Stub([0, 1, 2, 3])
    rms = m.fit(torch.from_numpy(X.reshape(300, 6, 1, 1)), torch.from_numpy(Z), chunk=64)
    assert rms < 1e-10
    assert np.allclose(m.W[:-1], Wt) and np.allclose(m.W[-1], bias)
    Xt = rng.standard_normal((4, 6))
    got = m.predict(torch.from_numpy(Xt.reshape(4, 6, 1, 1))).numpy()
    assert np.allclose(got, Xt @ Wt + bias, atol=1e-5)


def test_parallax_pupil_save_load_roundtrip(tmp_path):
    from fpm_amd import warmstart as ws
    m = ws.ParallaxPupil([0, 2, 5], iters=3)
    with pytest.raises(fpm_amd.FPMError):
        m.save(str(tmp_path / "unfitted.npz"))
    m.W = np.random.default_rng(1).standard_normal((5, 5))
    path = str(tmp_path / "pupil.npz")
    m.save(path)
    m2 = ws.ParallaxPupil.load(path)
    assert m2.planes == [0, 2, 5] and m2.iters == 3 and np.array_equal(m2.W, m.W)


def test_noalloc_gradient_rejects_a_call_beyond_its_scratch_before_device_work():
    """fpm_fused_gradient_noalloc checks the call's scratch need against the
    static array (FPM_NOALLOC_SCRATCH_BYTES) before touching a device: a
    2048^2 object needs 64 MiB of spectra alone."""
    L = fpm_amd.lib()
    buf = np.zeros(16, np.complex64)
    p = buf.ctypes.data
    rc = L.fpm_fused_gradient_noalloc(p, 2048, 2048, p, 512, 512, p, 26, p, 4.0, p, p, None, None)
    assert rc == -1 and "scratch" in L.fpm_last_error().decode()
    rc = L.fpm_fused_gradient_noalloc(p, 96, 96, p, 48, 48, p, 4, p, 2.0, p, p, None, None)
    assert rc == -2 and "power-of-two" in L.fpm_last_error().decode()


def test_solve_FPM_passes_the_matlab_defaults(monkeypatch):
    """solve_FPM (codes_v2/main_run_FPM.m:77-107) forwards its name-value
    defaults pratio 8, epoch 20, batchSize 26 and [5e-4 0.999 1e-4] to the
    mex call, and overrides by keyword."""
    seen = {}

    def fake(*args, **kw):
        seen["args"], seen["kw"] = args, kw
        return args[0], args[1]

    monkeypatch.setattr(fpm_amd, "cuFPM_pure_v3", fake)
    O, P = np.zeros((64, 64), np.complex64), np.zeros((16, 16), np.complex64)
    fpm_amd.solve_FPM(O, P, np.zeros((16, 16, 3), np.float32), np.ones((2, 3), np.int32), np.ones((16, 16)))
    assert seen["args"][5:8] == (8, 20, 26) and seen["args"][8] == (0.0005, 0.999, 1e-4)
    assert seen["args"][4].dtype == np.float32
    fpm_amd.solve_FPM(O, P, None, None, np.ones((16, 16)), batchSize=16, learningRate=1e-3, graph=True)
    assert seen["args"][7] == 16 and seen["args"][8][0] == 1e-3 and seen["kw"] == {"graph": True}
