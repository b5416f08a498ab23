"""Long-signal form (config 5: 30 min at 44.1 kHz, 79.4 M samples, 1.8 M
one-millisecond frames) against WORLD's output for the same recording
(tests/golden/world_long.npz, written by make_golden_long.py from WORLD built
from the reference).  The bar is the plain path's: identical voiced/unvoiced
decisions and relative F0 error <= 1e-6 (the golden F0 is stored as float32,
whose rounding, 6e-8, is far inside the bar)."""
import hashlib

import numpy as np
import pytest
from conftest import GOLDEN

from harvest_cuda.signals import config5_recording

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def recording():
    x, fs = config5_recording(0)
    return x, fs


def test_config5_recording_matches_world(gpu, recording):
    g = np.load(GOLDEN / "world_long.npz")
    x, fs = recording
    assert hashlib.sha256(x.tobytes()).hexdigest() == str(g["sha"])
    f0, t = gpu.harvest(x, fs)
    ref = g["f0"].astype(np.float64)
    assert f0.shape == ref.shape == (360001,)
    np.testing.assert_array_equal(t, np.arange(len(t)) * 5.0 / 1000.0)
    v = ref > 0
    assert np.array_equal(f0 > 0, v), np.count_nonzero((f0 > 0) != v)
    assert np.max(np.abs(f0[v] - ref[v]) / ref[v]) <= 1e-6
    assert v.mean() > 0.3


def test_stream_of_long_recordings(gpu, recording):
    from harvest_cuda import harvest_stream
    x, fs = recording
    half = np.ascontiguousarray(x[: len(x) // 2])
    outs = list(harvest_stream(iter([x, half]), fs))
    assert len(outs) == 2
    np.testing.assert_array_equal(outs[0][0], gpu.harvest(x, fs)[0])
    np.testing.assert_array_equal(outs[1][0], gpu.harvest(half, fs)[0])
