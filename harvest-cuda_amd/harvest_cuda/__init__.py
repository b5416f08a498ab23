"""MI355X-native (gfx950, HIP) WORLD Harvest F0 estimator.

Drop-in for ``harvest_cuda.harvest`` / ``pyworld.harvest``::

    f0, t = harvest(x, fs, f0_floor=71.0, f0_ceil=800.0, frame_period=5.0, device=0)

There is no CPU fallback: without a supported GPU the call raises RuntimeError.
"""
# torch and this package must share one HIP runtime.  torch ships its own
# libamdhip64 with the same SONAME (libamdhip64.so.7) as ROCm's; loading torch
# first makes the backend bind to that copy instead of mapping a second one.
try:
    import torch as _torch  # noqa: F401
except ImportError:  # pragma: no cover - torch is optional for the NumPy API
    _torch = None

from ._core import clear_cache, cuda_available, harvest, harvest_batch  # noqa: E402
from .stream import harvest_stream  # noqa: E402
from . import tuning as _tuning  # noqa: E402

_tuning.apply()  # the autotuned table (harvest_cuda/tuning.json), if present

__all__ = ["harvest", "harvest_batch", "harvest_stream", "cuda_available", "clear_cache"]
__version__ = "0.1.0"
