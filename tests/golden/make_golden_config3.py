"""Writes tests/golden/world_config3.npz: WORLD Harvest's F0 (5 ms frames,
float64) of 24 utterances of the batched workload (512 utterances at 24 kHz,
lengths log-uniform over 0.5-30 s, seed 1003), computed by WORLD compiled
from the reference checkout (oracle/_ref).  The 24 are the utterances of
length ranks round(linspace(0, 511, 24)), so they span the whole length
range.  The SHA-256 of each input is stored so that a change of the signal
generator is detected.  Run from the repo root:
    python tests/golden/make_golden_config3.py
"""
import hashlib
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "harvest-cuda_amd"))
sys.path.insert(0, str(ROOT))

from harvest_cuda.signals import config3  # noqa: E402
from oracle import oracle  # noqa: E402

SUBSET = 24


def subset_indices(lengths) -> np.ndarray:
    """Utterances at evenly spaced length ranks, shortest to longest."""
    order = np.argsort(np.asarray(lengths), kind="stable")
    ranks = np.round(np.linspace(0, len(order) - 1, SUBSET)).astype(int)
    return order[ranks]


def main():
    if not oracle.ref_available():
        oracle.build(ref=True)
    xs, fs = config3()
    idx = subset_indices([len(x) for x in xs])
    t0 = time.perf_counter()
    with ThreadPoolExecutor(8) as pool:  # the WORLD call releases the GIL
        f0s = list(pool.map(lambda i: oracle.ref_harvest(xs[i], fs)[0], idx))
    print(f"WORLD on {len(idx)} utterances: {time.perf_counter() - t0:.1f} s", flush=True)
    out = {"index": idx, "fs": fs}
    for i, f0 in zip(idx, f0s):
        out[f"{i}_f0"] = f0
        out[f"{i}_sha"] = np.array(hashlib.sha256(xs[i].tobytes()).hexdigest())
    path = Path(__file__).parent / "world_config3.npz"
    np.savez_compressed(path, **out)
    print(path.name, path.stat().st_size)


if __name__ == "__main__":
    main()
