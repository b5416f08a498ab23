"""Long recordings, one after another (e.g. 30-minute recordings at 44.1 kHz).

A 30-minute recording at 44.1 kHz (79.4 M samples, 1.8 M one-millisecond
frames) is analysed in one call: the filterbank runs as direct FIR tiles over
band groups whose crossing slabs stay within 4 GiB, so no global-size spectra
are ever materialised and a recording needs about 15 GB of device memory.
``harvest_stream`` walks an iterable of recordings; a helper thread fetches
(or loads, or synthesises) the next recording while the current one is being
analysed, which the GIL release of the analysis makes concurrent.
"""
from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor
from typing import Iterable, Iterator

import numpy as np

from . import _core

__all__ = ["harvest_stream"]

_END = object()


def harvest_stream(recordings: Iterable[np.ndarray], fs: int, f0_floor: float = 71.0,
                   f0_ceil: float = 800.0, frame_period: float = 5.0, *,
                   device: int = 0) -> Iterator[tuple[np.ndarray, np.ndarray]]:
    """Yield (f0, t) for each recording of the iterable, in order."""
    it = iter(recordings)
    with ThreadPoolExecutor(max_workers=1) as pool:
        pending = pool.submit(next, it, _END)
        while True:
            x = pending.result()
            if x is _END:
                return
            pending = pool.submit(next, it, _END)
            yield _core.harvest(x, fs, f0_floor, f0_ceil, frame_period, device=device)
