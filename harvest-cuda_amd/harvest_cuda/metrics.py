"""Accuracy metrics of F0 tracks.

``compare`` reports the figures of the reference's benchmark
(benchmarks/benchmark.py:20-37; same keys and error texts): the share of
frames whose voicing differs, and the pitch error in cents,
1200 |log2(f_actual / f_reference)|, over frames voiced in both tracks
(its 99th percentile with linear interpolation, its maximum, and how many
frames are off by more than half an octave).  ``parity`` is the acceptance
bar of this build: identical voicing and the largest relative error.
"""
from __future__ import annotations

import numpy as np

_HALF_OCTAVE_CENTS = 600.0


def _validated(label: str, track) -> np.ndarray:
    f = np.asarray(track, dtype=np.float64)
    if np.any(~np.isfinite(f) | (f < 0)):
        raise ValueError(f"{label} F0 must contain only finite, nonnegative values")
    return f


def compare(reference: np.ndarray, actual: np.ndarray) -> dict:
    """Voicing agreement and pitch error in cents of `actual` against `reference`."""
    if np.shape(reference) != np.shape(actual):
        raise ValueError(f"Shape mismatch: {np.shape(reference)} versus {np.shape(actual)}")
    ref = _validated("reference", reference)
    act = _validated("actual", actual)
    ref_on, act_on = ref > 0, act > 0
    joint = ref_on & act_on
    frames = ref.size
    # Cents between the two voiced values of each jointly voiced frame.
    cents = 1200.0 * np.abs(np.log2(act[joint]) - np.log2(ref[joint]))
    summary = {
        "frames": int(frames),
        "reference_voiced_frames": int(np.count_nonzero(ref_on)),
        "joint_voiced_frames": int(np.count_nonzero(joint)),
        "vuv_disagreement": float(np.count_nonzero(ref_on != act_on)) / frames if frames else 0.0,
        "cents_p99": None,
        "cents_max": None,
        "over_600_cents": int(np.count_nonzero(cents > _HALF_OCTAVE_CENTS)),
    }
    if cents.size:
        summary["cents_p99"] = float(np.percentile(cents, 99.0))
        summary["cents_max"] = float(np.max(cents))
    return summary


def parity(expected: np.ndarray, actual: np.ndarray) -> dict:
    """Identical voicing and the largest relative F0 error on frames voiced in
    both tracks."""
    v = expected > 0
    both = v & (actual > 0)
    rel = float(np.max(np.abs(actual[both] - expected[both]) / expected[both])) if both.any() else 0.0
    return {"vuv_identical": bool(np.array_equal(v, actual > 0)),
            "vuv_mismatches": int(np.count_nonzero(v != (actual > 0))),
            "max_rel_err": rel}
