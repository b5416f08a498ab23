"""The two refinement paths (needs an MI355X).

At working rates that are multiples of 1000 Hz (every 8 kHz working rate)
the 1 ms frames start at whole samples and the refinement runs the bin-table
path (refine_table.hip); other rates run the per-frame kernel.  Both restate
GetMeanF0 and FixF0 (harvest.cpp:507-584); they evaluate the same DFT bins
with different operation orders, so their F0 agree to rounding and each
meets the parity bar against the oracle."""
import numpy as np
import pytest
from ref_signals import reference_signal

pytestmark = pytest.mark.gpu
RTOL = 1e-6


def assert_parity(expected, actual, rtol):
    v = expected > 0
    assert np.array_equal(v, actual > 0), f"{np.count_nonzero(v != (actual > 0))} V/UV mismatches"
    if v.any():
        rel = np.max(np.abs(actual[v] - expected[v]) / expected[v])
        assert rel <= rtol, rel


@pytest.mark.parametrize("fs,table", [(8000, True), (16000, True), (24000, True), (48000, True),
                                      (96000, True), (22050, False), (44100, False)])
def test_path_follows_working_rate(gpu, fs, table):
    x = reference_signal("speechlike", seconds=0.9, fs=fs)
    out, tm = gpu.harvest_batch([x], fs, timings=True)
    assert (tm["work_table_pairs"] > 0) == table, tm["work_table_pairs"]
    assert tm["work_bin_samples"] > 0


@pytest.mark.parametrize("kind", ["speechlike", "noise", "high", "low"])
@pytest.mark.parametrize("fs", [16000, 48000])
def test_bin_table_matches_per_frame_path(gpu, oracle, fs, kind):
    from harvest_cuda import tuning
    saved = tuning.current()
    xs = [reference_signal(kind, seconds=s, fs=fs) for s in (0.6, 3.3)]
    try:
        tuning.set_backend(refine_table=1)
        table, tm1 = gpu.harvest_batch(xs, fs, timings=True)
        tuning.set_backend(refine_table=0)
        frame, tm0 = gpu.harvest_batch(xs, fs, timings=True)
    finally:
        tuning.apply(saved)
    assert tm1["work_table_pairs"] > 0 and tm0["work_table_pairs"] == 0
    for x, (f1, t1), (f0, t0) in zip(xs, table, frame):
        np.testing.assert_array_equal(t1, t0)
        assert_parity(f0, f1, 1e-10)
        ef0, _ = oracle.harvest(x, fs)
        assert_parity(ef0, f1, RTOL)
        assert_parity(ef0, f0, RTOL)


def test_bin_table_batched_equals_plain(gpu):
    fs = 24000
    kinds = ["speechlike", "noise", "low", "high", "speechlike"]
    xs = [reference_signal(k, seconds=s, fs=fs) for k, s in zip(kinds, (2.2, 0.7, 5.1, 1.3, 9.6))]
    batched = gpu.harvest_batch(xs, fs)
    for x, (f0, t) in zip(xs, batched):
        pf0, pt = gpu.harvest(x, fs)
        np.testing.assert_array_equal(f0, pf0)
        np.testing.assert_array_equal(t, pt)


@pytest.mark.parametrize("kind", ["noise", "speechlike"])
def test_frames_beyond_the_pair_limit(gpu, oracle, kind):
    """A frame's slots can span more half windows than the bin-table path
    lists per frame (32); such frames compute their bins on their own.  At
    16 kHz with a 50-1500 Hz search range, 2 s of these signals hold 10-12
    such frames (counted on the oracle's slots)."""
    from harvest_cuda import tuning
    saved = tuning.current()
    fs, floor, ceil = 16000, 50.0, 1500.0
    x = reference_signal(kind, seconds=2.0, fs=fs)
    try:
        tuning.set_backend(refine_table=1)
        (table,), tm1 = gpu.harvest_batch([x], fs, floor, ceil, timings=True)
        tuning.set_backend(refine_table=0)
        (frame,), tm0 = gpu.harvest_batch([x], fs, floor, ceil, timings=True)
    finally:
        tuning.apply(saved)
    assert tm1["work_table_pairs"] > 0 and tm0["work_table_pairs"] == 0
    ef0, _ = oracle.harvest(x, fs, floor, ceil)
    assert_parity(ef0, table[0], RTOL)
    assert_parity(frame[0], table[0], 1e-10)


def test_dft_grid_leaves_results_unchanged(gpu):
    """The DFT kernel's persistent grid (refine_groups) only changes which
    workgroup computes a pair: results are bitwise the same at any size."""
    from harvest_cuda import tuning
    saved = tuning.current()
    fs = 16000
    xs = [reference_signal(k, seconds=s, fs=fs) for k, s in (("speechlike", 7.5), ("noise", 2.0))]
    try:
        tuning.set_backend(refine_groups=0)
        want = gpu.harvest_batch(xs, fs)
        for groups in (64, 333, 4096):
            tuning.set_backend(refine_groups=groups)
            got = gpu.harvest_batch(xs, fs)
            for (f0, _), (wf0, _) in zip(got, want):
                np.testing.assert_array_equal(f0, wf0, err_msg=f"refine_groups {groups}")
    finally:
        tuning.apply(saved)
