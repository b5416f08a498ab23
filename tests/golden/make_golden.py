"""Writes the golden fixtures of tests/golden/ from WORLD itself.

WORLD Harvest is compiled from the reference checkout into
oracle/_ref/libworld_ref.so (oracle/Makefile, target `ref`); this script calls
it through oracle/oracle.py and stores inputs' fingerprints and outputs:

  world_harvest.npz   f0/t of WORLD Harvest for the reference's test signal
                      kinds at several fs and frame periods
  world_stages.npz    WORLD's intermediate arrays for one signal (decimated
                      waveform, one band's filtered signal, raw candidates,
                      official+overlapped candidates, refined, pruned,
                      FixF0Contour, SmoothF0Contour)
  world_contour.npz   FixF0Contour / SmoothF0Contour on random candidate
                      matrices
  world_edges.npz     edge inputs: short silences, all decimation ratios,
                      low floors, high f0, Nyquist-bin excitation

Inputs are regenerated deterministically by harvest_cuda.signals and the
generators below; the SHA-256 of each input is stored to detect drift.
Run from the repo root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "harvest-cuda_amd"))
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from ref_signals import reference_signal  # noqa: E402
from oracle import oracle  # noqa: E402

OUT = Path(__file__).resolve().parent
HARVEST_CASES = [(fs, kind, 5.0) for fs in (8000, 16000, 24000, 44100, 48000)
                 for kind in ("speechlike", "low", "high", "noise")] + \
                [(16000, "speechlike", p) for p in (1.0, 2.5, 10.0)] + \
                [(44100, "speechlike", 7.3)]
STAGE_CASE = (16000, "speechlike", 0.35)


def sha(x: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(x, dtype=np.float64).tobytes()).hexdigest()


def random_candidates(trial: int):
    """Random candidate matrices in the style of the reference's contour test
    (tests/cpp/test_contour.cpp:116-145): tracks with crossable and
    non-crossable gaps."""
    rng = np.random.default_rng(9000 + trial)
    pitches = np.array([80, 90, 120, 150, 180, 250, 400, 600.0])
    frames, cols = 2 + int(rng.integers(1500)), 8
    f0 = np.zeros((frames, cols))
    sc = np.zeros((frames, cols))
    rows = np.arange(frames)
    for col in range(cols):
        period, gap, phase = 10 + int(rng.integers(150)), int(rng.integers(8)), int(rng.integers(160))
        voiced = (trial % 4 == 0) | (((rows + phase) % period) >= gap)
        jitter = 1 + 0.02 * rng.standard_normal(frames) * (trial % 5 == 1)
        val = pitches[col % 4 if trial % 3 == 0 else col] * jitter
        f0[:, col] = np.where(voiced, val, 0)
        sc[:, col] = np.where(voiced, 1 + rng.integers(10, size=frames), 0)
    if trial % 10 == 0:
        f0[:] = 0
        sc[:] = 0
    return f0, sc


def edge_cases():
    rng = np.random.default_rng(456)
    cases = []
    for n in (6, 24, 240):
        cases.append((f"silence_{n}", np.zeros(n), 24000, 71.0, 800.0))
    for fs in (16000, 24000, 32000, 40000, 48000, 56000, 64000, 72000, 80000, 88000, 96000,
               192000, 44100, 11025, 22050):
        x = reference_signal("speechlike", seconds=0.6, fs=fs)[:-1]
        cases.append((f"ratio_{fs}", x, fs, 71.0, 800.0))
    for fs, freq, lo, hi in ((8000, 16, 8, 90), (11025, 30, 10, 100), (192000, 75, 40, 250),
                             (8000, 2400, 1200, 3900)):
        seconds = 2.0 if freq < 40 else 0.8
        t = np.arange(round(seconds * fs)) / fs
        pitch = freq * (1 + 0.2 * np.sin(2 * np.pi * 0.7 * t))
        phase = 2 * np.pi * np.cumsum(pitch) / fs
        harm = min(6, int(fs / (2 * freq * 1.2)))
        x = sum(0.2 * np.sin(h * phase + 0.1 * h) / h for h in range(1, harm + 1))
        x = x + rng.normal(0, 0.0001, len(t))
        cases.append((f"window_{fs}_{freq}", np.ascontiguousarray(x), fs, float(lo), float(hi)))
    for n in (24576, 24577, 24800):
        x = reference_signal("speechlike", seconds=n / 8000, fs=8000)
        i = np.arange(len(x))
        x = x + 0.01 * ((1 - 2 * (i % 2)) + np.cos(2 * np.pi * (0.5 - 1 / 32768) * i))
        cases.append((f"nyquist_{n}", np.ascontiguousarray(x), 8000, 71.0, 800.0))
    return cases


def main():
    if not oracle.ref_available():
        oracle.build(ref=True)
    out = {}
    for fs, kind, period in HARVEST_CASES:
        x = reference_signal(kind, seconds=0.35, fs=fs)
        f0, t = oracle.ref_harvest(x, fs, period=period)
        key = f"{fs}_{kind}_{period}"
        out[key + "_f0"] = f0
        out[key + "_t"] = t
        out[key + "_sha"] = np.array(sha(x))
    np.savez_compressed(OUT / "world_harvest.npz", **out)

    fs, kind, seconds = STAGE_CASE
    x = reference_signal(kind, seconds=seconds, fs=fs)
    p = oracle.params(len(x), fs)
    y = oracle.ref_waveform(x, fs)
    band = 60
    filtered = oracle.ref_filtered_band(x, fs, band)
    raw = oracle.ref_raw_candidates(x, fs)
    cand, count = oracle.ref_official(raw, p)
    ncand = 7 * count
    rf0, rsc = oracle.ref_refine(y, p, cand, ncand)
    pf0, psc = oracle.ref_prune(rf0, rsc, ncand)
    contour = oracle.ref_fix_contour(pf0, psc, ncand)
    smooth = oracle.ref_smooth(contour)
    np.savez_compressed(OUT / "world_stages.npz", sha=np.array(sha(x)), fs=fs, seconds=seconds,
                        y=y, band=band, filtered=filtered, raw=raw, cand=cand, count=count,
                        refined_f0=rf0, refined_score=rsc, pruned_f0=pf0, pruned_score=psc,
                        contour=contour, smooth=smooth)

    out = {}
    for trial in range(40):
        f0, sc = random_candidates(trial)
        c = oracle.ref_fix_contour(f0, sc, f0.shape[1])
        out[f"{trial}_contour"] = c
        out[f"{trial}_smooth"] = oracle.ref_smooth(c)
    np.savez_compressed(OUT / "world_contour.npz", **out)

    out = {}
    for name, x, fs, lo, hi in edge_cases():
        f0, t = oracle.ref_harvest(x, fs, lo, hi)
        out[name + "_f0"] = f0
        out[name + "_t"] = t
        out[name + "_sha"] = np.array(sha(x))
    np.savez_compressed(OUT / "world_edges.npz", **out)
    for f in sorted(OUT.glob("*.npz")):
        print(f.name, f.stat().st_size)


if __name__ == "__main__":
    main()
