"""The batched workload at full size (needs an MI355X): 512 utterances at
24 kHz, lengths log-uniform over 0.5-30 s (seed 1003), in one batched call.

  * every utterance's F0 and time axis equal its plain call's, bit for bit;
  * 24 utterances spanning the length range match WORLD Harvest
    (tests/golden/world_config3.npz, written by make_golden_config3.py from
    WORLD built from the reference): identical voiced/unvoiced decisions and
    |f0 - f0_WORLD| / f0_WORLD <= 1e-6 on voiced frames, with the reference
    benchmark's accuracy figures (benchmarks/benchmark.py:20-37) reported
    beside the bar.
"""
import hashlib

import numpy as np
import pytest
from conftest import GOLDEN

from harvest_cuda.metrics import compare
from harvest_cuda.signals import config3

pytestmark = pytest.mark.gpu
RTOL = 1e-6


@pytest.fixture(scope="module")
def workload():
    return config3()


@pytest.fixture(scope="module")
def batched(gpu, workload):
    xs, fs = workload
    return gpu.harvest_batch(xs, fs)


def test_config3_batched_equals_plain_bitwise(gpu, workload, batched):
    xs, fs = workload
    assert len(xs) == len(batched) == 512
    assert abs(sum(len(x) for x in xs) / fs - 3568.31) < 0.01
    for u, (x, (bf0, bt)) in enumerate(zip(xs, batched)):
        f0, t = gpu.harvest(x, fs)
        np.testing.assert_array_equal(bt, t, err_msg=f"utterance {u}")
        np.testing.assert_array_equal(bf0, f0, err_msg=f"utterance {u}")


def test_config3_subset_matches_world(workload, batched):
    xs, fs = workload
    g = np.load(GOLDEN / "world_config3.npz")
    idx = [int(i) for i in g["index"]]
    assert len(idx) == 24 and int(g["fs"]) == fs
    lengths = sorted(len(xs[i]) for i in idx)
    assert lengths[0] < 0.6 * fs and lengths[-1] > 29 * fs  # spans 0.5-30 s
    worst = 0.0
    for i in idx:
        x = xs[i]
        assert hashlib.sha256(x.tobytes()).hexdigest() == str(g[f"{i}_sha"]), i
        ref = g[f"{i}_f0"]
        f0, t = batched[i]
        assert f0.shape == ref.shape, i
        np.testing.assert_array_equal(t, np.arange(len(t)) * 5.0 / 1000.0)
        v = ref > 0
        assert np.array_equal(f0 > 0, v), (i, int(np.count_nonzero((f0 > 0) != v)))
        if v.any():
            rel = float(np.max(np.abs(f0[v] - ref[v]) / ref[v]))
            assert rel <= RTOL, (i, rel)
            worst = max(worst, rel)
        acc = compare(ref, f0)
        assert acc["vuv_disagreement"] == 0.0 and acc["over_600_cents"] == 0, (i, acc)
    print(f"config-3 subset vs WORLD: {len(idx)} utterances, max rel err {worst:.3g}")


def test_config3_timings_leave_results_unchanged(gpu, workload, batched):
    """The pipelined batch with per-kernel timings (hipEvents around every
    launch on both workspaces' streams) returns the same F0 and sums the
    groups' kernel times and work counters."""
    xs, fs = workload
    out, tm = gpu.harvest_batch(xs, fs, timings=True)
    for (f0, t), (bf0, bt) in zip(out, batched):
        np.testing.assert_array_equal(f0, bf0)
        np.testing.assert_array_equal(t, bt)
    assert tm["groups"] > 1
    for key in ("kernel_refine", "kernel_filterbank", "kernel_interp", "work_bin_samples"):
        assert tm[key] > 0, key
    assert tm["kernel_refine"] < tm["preprocess"] + tm["filter_events"] + tm["refine"]


def test_batch_sizes_its_buffers_once(gpu, workload, batched):
    """The pipelined batch reserves each workspace's buffers for its largest
    bucket before submitting: on new workspaces no device buffer is freed to
    grow (the driver clears freed device memory in the background, which
    slowed the calls that followed), and once the pinned candidate buffers
    have adapted to the batch's candidate count (after the first call),
    repeating the call allocates nothing; the F0 stay the same."""
    xs, fs = workload
    gpu.clear_cache()
    _, regrowths = gpu._allocation_stats()
    first = gpu.harvest_batch(xs, fs)
    assert gpu._allocation_stats()[1] == regrowths
    second = gpu.harvest_batch(xs, fs)
    stats = gpu._allocation_stats()
    assert stats[1] == regrowths
    third = gpu.harvest_batch(xs, fs)
    assert gpu._allocation_stats() == stats
    for (f0, _), (f1, _), (f2, _), (bf0, _) in zip(first, second, third, batched):
        np.testing.assert_array_equal(f0, bf0)
        np.testing.assert_array_equal(f1, bf0)
        np.testing.assert_array_equal(f2, bf0)
