"""Test signals of the reference's test suite and benchmark.

``reference_signal`` reproduces the reference's test input generator
(benchmarks/signals.py:5-24: kinds speechlike/low/high/noise/silence, seed
20261110).  The inputs must be the same samples for the parity cases to mean
what the reference's cases mean, so the seed, the formula and its order of
operations follow that generator.  Test infrastructure only.
"""
from __future__ import annotations

import numpy as np


def reference_signal(kind: str = "speechlike", seconds: float = 1.0, fs: int = 24000) -> np.ndarray:
    rng = np.random.default_rng(20261110)
    t = np.arange(round(seconds * fs)) / fs
    if kind == "silence":
        return np.zeros(len(t), dtype=np.float64)
    if kind == "noise":
        return rng.normal(0, 0.05, len(t))
    if kind == "low":
        freq = 80 + 5 * np.sin(2 * np.pi * 2 * t)
    elif kind == "high":
        freq = 650 + 30 * np.sin(2 * np.pi * 3 * t)
    else:
        freq = 160 + 45 * np.sin(2 * np.pi * 0.7 * t) + 3 * np.sin(2 * np.pi * 5 * t)
    phase = 2 * np.pi * np.cumsum(freq) / fs
    x = sum(np.sin(h * phase + h * 0.1) / h ** 1.3 for h in range(1, 11))
    env = np.minimum(1, t / 0.02) * np.minimum(1, (seconds - t) / 0.02)
    if kind == "speechlike":
        env = env * np.clip(8 * np.sin(2 * np.pi * 1.8 * t), 0, 1)
    return np.ascontiguousarray(0.2 * x * env + rng.normal(0, 0.0002, len(t)))
