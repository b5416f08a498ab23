"""Writes tests/golden/world_long.npz: WORLD Harvest's F0 (5 ms frames) of
the config-5 recording 0 (30 min at 44.1 kHz, seed 1005), computed by WORLD
compiled from the reference checkout (oracle/_ref), plus the SHA-256 of the
input so that a change of the signal generator is detected.  F0 is stored
as float32 (relative rounding 6e-8, far below the 1e-6 parity bar; zeros
stay exact).  Needs about 25 GB of RAM and several minutes of one core."""
import hashlib
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "harvest-cuda_amd"))
sys.path.insert(0, str(ROOT))

from harvest_cuda.signals import config5_recording  # noqa: E402
from oracle import oracle  # noqa: E402


def main():
    x, fs = config5_recording(0)
    sha = hashlib.sha256(x.tobytes()).hexdigest()
    print(f"recording 0: {len(x)} samples, sha {sha[:16]}", flush=True)
    t0 = time.perf_counter()
    f0, t = oracle.ref_harvest(x, fs)
    seconds = time.perf_counter() - t0
    print(f"WORLD: {seconds:.1f} s, {np.count_nonzero(f0)} voiced of {len(f0)}", flush=True)
    np.savez_compressed(Path(__file__).parent / "world_long.npz", f0=f0.astype(np.float32),
                        sha=sha, world_seconds=seconds, samples=len(x), fs=fs)


if __name__ == "__main__":
    main()
