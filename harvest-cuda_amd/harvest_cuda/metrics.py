"""Accuracy metrics of the benchmark method (reference benchmarks/benchmark.py:20-37)."""
from __future__ import annotations

import numpy as np


def compare(reference: np.ndarray, actual: np.ndarray) -> dict:
    """Voicing agreement and pitch error in cents between two F0 tracks."""
    if reference.shape != actual.shape:
        raise ValueError(f"Shape mismatch: {reference.shape} versus {actual.shape}")
    for name, f0 in (("reference", reference), ("actual", actual)):
        if not np.all(np.isfinite(f0)) or np.any(f0 < 0):
            raise ValueError(f"{name} F0 must contain only finite, nonnegative values")
    rv, av = reference > 0, actual > 0
    both = rv & av
    cents = np.abs(1200 * np.log2(actual[both] / reference[both]))
    return {
        "frames": len(reference),
        "reference_voiced_frames": int(rv.sum()),
        "joint_voiced_frames": int(both.sum()),
        "vuv_disagreement": float(np.mean(rv != av)),
        "cents_p99": float(np.quantile(cents, 0.99)) if len(cents) else None,
        "cents_max": float(cents.max()) if len(cents) else None,
        "over_600_cents": int((cents > 600).sum()),
    }


def parity(expected: np.ndarray, actual: np.ndarray) -> dict:
    """The acceptance bar of this build: identical voicing and the largest
    relative F0 error on frames voiced in both."""
    v = expected > 0
    both = v & (actual > 0)
    rel = float(np.max(np.abs(actual[both] - expected[both]) / expected[both])) if both.any() else 0.0
    return {"vuv_identical": bool(np.array_equal(v, actual > 0)),
            "vuv_mismatches": int(np.count_nonzero(v != (actual > 0))),
            "max_rel_err": rel}
