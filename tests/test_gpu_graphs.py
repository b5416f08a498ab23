"""HIP graph capture of the launch sequence (needs an MI355X).

With graphs on, each analysis' device work is captured once per plan shape
and buffer set and replayed afterwards.  Every replay must equal the
launch-by-launch result bit for bit, with eager work (an analysis with
timings, which runs launch by launch, and plain calls of other shapes) and
synchronisations between the replays; the backend's counters show that
replays happened."""
import numpy as np
import pytest
from ref_signals import reference_signal

pytestmark = pytest.mark.gpu


def test_graph_replays_equal_eager(gpu):
    from harvest_cuda import tuning
    saved = tuning.current()
    fs = 16000
    kinds = ["speechlike", "low", "noise", "speechlike", "high", "speechlike"]
    secs = [41.0, 33.5, 27.0, 45.25, 30.0, 38.0]
    xs = [reference_signal(k, seconds=s, fs=fs) for k, s in zip(kinds, secs)]
    short = [reference_signal("speechlike", seconds=s, fs=fs) for s in (0.8, 2.1, 5.3)]
    try:
        tuning.set_backend(graphs=0)
        want_batch = gpu.harvest_batch(xs, fs)
        want_short = [gpu.harvest(x, fs) for x in short]
        tuning.set_backend(graphs=1)
        r0, c0 = gpu._graph_stats(0)
        for rep in range(4):
            got = gpu.harvest_batch(xs, fs)
            for (f0, t), (wf0, wt) in zip(got, want_batch):
                np.testing.assert_array_equal(f0, wf0, err_msg=f"replay {rep}")
                np.testing.assert_array_equal(t, wt)
            out, tm = gpu.harvest_batch(xs[:2], fs, timings=True)  # eager, launch by launch
            np.testing.assert_array_equal(out[0][0], want_batch[0][0])
            for x, (wf0, _) in zip(short, want_short):
                np.testing.assert_array_equal(gpu.harvest(x, fs)[0], wf0)
        r1, c1 = gpu._graph_stats(0)
        assert c1 > c0, "nothing was captured"
        assert r1 - r0 >= 3 * 2, "the repeated batch did not replay its graphs"
    finally:
        tuning.apply(saved)
