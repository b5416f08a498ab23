"""The reference engine's binding of the C ABI (needs an MI355X).

The reference's analyze() (src/engine.cpp:254-286) decimates with
gpu_decimate, removes the mean on the host with a sequential sum, and calls
gpu_analyze with WORLD's sizes and band boundaries.  These tests drive the
replacements exactly that way through ctypes: hv_decimate, the host mean
removal, hv_analyze and hv_fetch_candidates, and check the candidates
against the oracle's stages (CSR offsets exact, refined f0 within 1e-9 and
scores within 1e-6 relative, best F0 at the parity bar), plus
hv_analyze's argument checks."""
import ctypes
from pathlib import Path

import numpy as np
import pytest
from ref_signals import reference_signal

pytestmark = pytest.mark.gpu
P, D, I64 = ctypes.c_void_p, ctypes.c_double, ctypes.c_int64


@pytest.fixture(scope="module")
def lib(gpu):
    lib = ctypes.CDLL(str(Path(gpu.__file__).parent / "libharvest_hip.so"))
    lib.hv_last_error.restype = ctypes.c_char_p
    lib.hv_workspace_create.argtypes = [ctypes.c_int, P]
    lib.hv_workspace_destroy.argtypes = [P]
    lib.hv_decimate.argtypes = [P, P, I64, ctypes.c_int, P, I64]
    lib.hv_analyze.argtypes = [P, P, I64, ctypes.c_int, D, P, ctypes.c_int, ctypes.c_int, D, D,
                               ctypes.c_int, P, P, P, P]
    lib.hv_fetch_candidates.argtypes = [P, P, P]
    return lib


@pytest.fixture(scope="module")
def ws(lib):
    w = P()
    assert lib.hv_workspace_create(0, ctypes.byref(w)) == 0, lib.hv_last_error()
    yield w
    lib.hv_workspace_destroy(w)


def world_setup(length, fs, floor, ceil):
    """analyze()'s sizes (src/engine.cpp:257-266): bands and boundaries,
    ratio, working rate, FFT size, 1 ms frames and candidate capacity."""
    bands = 1 + int(np.log(ceil * 1.1 / (floor * 0.9)) / 0.69314718055994529 * 40)
    bounds = np.array([floor * 0.9 * 2.0 ** ((i + 1) / 40.0) for i in range(bands)])
    ratio = min(max(int(fs / 8000.0 + 0.5), 1), 12)
    y_length = int(np.ceil(length / ratio))
    fs_w = fs / ratio
    sample = y_length + 5 + 2 * int(2 * fs_w / bounds[0])
    fft = int(2.0 ** (int(np.log(sample) / 0.69314718055994529) + 1))
    frames = int(1000.0 * length / fs) + 1
    capacity = int(bands / 10.0 + 0.5)
    return bounds, ratio, y_length, fs_w, fft, frames, capacity


@pytest.mark.parametrize("fs,floor,ceil", [(16000, 71.0, 800.0), (44100, 71.0, 800.0),
                                            (48000, 100.0, 600.0), (8000, 71.0, 800.0)])
def test_reference_binding_matches_oracle(lib, ws, oracle, fs, floor, ceil):
    x = reference_signal("speechlike", seconds=1.3, fs=fs)
    bounds, ratio, n, fs_w, fft, frames, cap = world_setup(len(x), fs, floor, ceil)
    p = oracle.params(len(x), fs, floor, ceil)
    assert (p.bands, p.ratio, p.y_length, p.fft_size, p.frames, p.capacity) == \
        (len(bounds), ratio, n, fft, frames, cap)
    np.testing.assert_array_equal(bounds, oracle.boundaries(p, floor))
    y = np.zeros(n)
    assert lib.hv_decimate(ws, x.ctypes.data, len(x), ratio, y.ctypes.data, n) == 0, \
        lib.hv_last_error()
    ey = oracle.decimate(x, fs, floor, ceil, remove_mean=False)
    np.testing.assert_allclose(y, ey, rtol=0, atol=1e-12 * np.max(np.abs(ey)))
    y -= np.cumsum(y)[-1] / n  # the host's sequential mean (src/engine.cpp:278-281)
    off = np.zeros(frames + 1, np.int32)
    best = np.zeros(frames)
    nnz = I64()
    assert lib.hv_analyze(ws, y.ctypes.data, n, fft, fs_w, bounds.ctypes.data, len(bounds),
                          frames, floor, ceil, cap, off.ctypes.data, best.ctypes.data,
                          ctypes.byref(nnz), None) == 0, lib.hv_last_error()
    f0, sc = np.zeros(max(nnz.value, 1)), np.zeros(max(nnz.value, 1))
    assert lib.hv_fetch_candidates(ws, f0.ctypes.data, sc.ctypes.data) == 0
    f0, sc = f0[:nnz.value], sc[:nnz.value]

    st = oracle.stages(x, fs, floor, ceil)
    ref, refs = st["f0"][:, :st["ncand"]], st["score"][:, :st["ncand"]]
    nz = ref != 0
    assert nnz.value == off[-1] == nz.sum() > 0
    np.testing.assert_array_equal(np.diff(off), nz.sum(1))
    np.testing.assert_allclose(f0, ref[nz], rtol=1e-9)
    np.testing.assert_allclose(sc, refs[nz], rtol=1e-6)
    v = st["best"] > 0
    assert np.array_equal(best > 0, v)
    assert np.max(np.abs(best[v] - st["best"][v]) / st["best"][v]) <= 1e-6


def test_analyze_rejects_invalid_arguments(lib, ws):
    fs = 16000
    x = reference_signal("speechlike", seconds=0.5, fs=fs)
    bounds, ratio, n, fs_w, fft, frames, cap = world_setup(len(x), fs, 71.0, 800.0)
    y = np.ascontiguousarray(x[::ratio][:n])
    off = np.zeros(frames + 1, np.int32)
    best = np.zeros(frames)
    nnz = I64()

    def call(floor=71.0, ceil=800.0, b=bounds, fft_size=fft, fsw=fs_w):
        return lib.hv_analyze(ws, y.ctypes.data, n, fft_size, fsw, b.ctypes.data, len(b),
                              frames, floor, ceil, cap, off.ctypes.data, best.ctypes.data,
                              ctypes.byref(nnz), None)

    assert call() == 0
    for kw, text in (({"floor": 0.0}, b"f0_floor"), ({"floor": -5.0}, b"f0_floor"),
                     ({"ceil": float("inf")}, b"f0_floor"), ({"ceil": 60.0}, b"f0_floor"),
                     ({"ceil": 4000.0}, b"Nyquist"), ({"fft_size": fft - 1}, b"fft_size"),
                     ({"fsw": float("nan")}, b"fs_work")):
        assert call(**kw) == 1, kw  # HV_EINVAL
        assert text in lib.hv_last_error(), (kw, lib.hv_last_error())
    bad = bounds.copy()
    bad[3] = np.nan
    assert call(b=bad) == 1 and b"boundaries" in lib.hv_last_error()
    bad[3] = -1.0
    assert call(b=bad) == 1 and b"boundaries" in lib.hv_last_error()
    assert call() == 0  # the workspace is still usable
