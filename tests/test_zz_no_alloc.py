"""The allocation-free decimation entry point (needs an MI355X).

hv_decimate_device_noalloc keeps its plan and scan arrays in a static device
array instead of workspace buffers.  Its output must equal hv_decimate's bit
for bit; two calls in a row (the second smaller, so that the array holds the
first call's larger contents) check that nothing depends on what the array
held before, and an input beyond the array's size is rejected."""
import ctypes
from pathlib import Path

import numpy as np
import pytest
from ref_signals import reference_signal

pytestmark = pytest.mark.gpu
P, I64 = ctypes.c_void_p, ctypes.c_int64


@pytest.fixture(scope="module")
def lib(gpu):
    lib = ctypes.CDLL(str(Path(gpu.__file__).parent / "libharvest_hip.so"))
    lib.hv_last_error.restype = ctypes.c_char_p
    lib.hv_workspace_create.argtypes = [ctypes.c_int, P]
    lib.hv_workspace_destroy.argtypes = [P]
    lib.hv_decimate.argtypes = [P, P, I64, ctypes.c_int, P, I64]
    lib.hv_decimate_device_noalloc.argtypes = [ctypes.c_int, P, I64, ctypes.c_int, P, I64, P]
    return lib


def decimate_workspace(lib, x, ratio):
    n = -(-len(x) // ratio)
    y = np.zeros(n)
    ws = P()
    assert lib.hv_workspace_create(0, ctypes.byref(ws)) == 0, lib.hv_last_error()
    try:
        assert lib.hv_decimate(ws, x.ctypes.data, len(x), ratio, y.ctypes.data, n) == 0, \
            lib.hv_last_error()
    finally:
        lib.hv_workspace_destroy(ws)
    return y


def decimate_static(lib, x, ratio):
    import torch
    n = -(-len(x) // ratio)
    dx = torch.from_numpy(x).to("cuda:0")
    dy = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda:0")
    stream = torch.cuda.current_stream()
    stream.synchronize()
    assert lib.hv_decimate_device_noalloc(0, dx.data_ptr(), len(x), ratio, dy.data_ptr(), n,
                                          stream.cuda_stream) == 0, lib.hv_last_error()
    stream.synchronize()
    return dy.cpu().numpy()


@pytest.mark.parametrize("fs", [48000, 16000, 44100, 8000])
def test_noalloc_equals_workspace_decimation(lib, fs):
    ratio = min(max(int(fs / 8000.0 + 0.5), 1), 12)
    first = reference_signal("speechlike", seconds=9.0, fs=fs)
    second = np.ascontiguousarray(reference_signal("low", seconds=2.3, fs=fs)[:-7])
    for x in (first, second):  # two calls in a row, the second smaller
        want = decimate_workspace(lib, x, ratio)
        got = decimate_static(lib, x, ratio)
        assert got.shape == want.shape
        np.testing.assert_array_equal(got, want)


def test_noalloc_rejects_inputs_beyond_its_scratch(lib):
    import torch
    n = 8_000_000
    dx = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    dy = torch.zeros(n // 2, dtype=torch.float64, device="cuda:0")
    assert lib.hv_decimate_device_noalloc(0, dx.data_ptr(), n, 2, dy.data_ptr(), n // 2,
                                          None) == 1  # HV_EINVAL
    assert b"scratch" in lib.hv_last_error()
    assert lib.hv_decimate_device_noalloc(0, None, n, 2, dy.data_ptr(), n // 2, None) == 1
    assert lib.hv_decimate_device_noalloc(-1, dx.data_ptr(), n, 2, dy.data_ptr(), n // 2,
                                          None) == 1
    # Still usable after the rejections.
    x = reference_signal("speechlike", seconds=1.0, fs=16000)
    np.testing.assert_array_equal(decimate_static(lib, x, 2), decimate_workspace(lib, x, 2))
