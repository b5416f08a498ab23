"""Differentiable Harvest F0 for PyTorch (frozen-decision mode).

``harvest_f0(x, fs)`` takes a float64 tensor of shape (batch, samples) and
returns F0 of shape (batch, frames) at the output frame period.  The forward
pass is the ordinary batched GPU analysis; it also records every discrete
decision (which candidate each contour frame copies, the bridged gaps, the
voiced sections, the resampling indices).  With those held the output is a
smooth function of the waveforms, and ``backward`` evaluates its gradient
with a hand-written reverse pass: host transposes of resampling, smoothing and
gap bridging, then the GPU adjoint of the refinement DFT, mean removal and
decimation.  No reference counterpart: the reference is forward-only.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _core

__all__ = ["HarvestF0", "harvest_f0", "frozen_f0"]


def _rows(t: torch.Tensor) -> list[np.ndarray]:
    a = t.detach().to("cpu", torch.float64).numpy()
    return [np.ascontiguousarray(r) for r in a]


class HarvestF0(torch.autograd.Function):
    """torch.autograd.Function around the GPU analysis and its adjoint."""

    @staticmethod
    def forward(ctx, x, fs, f0_floor=71.0, f0_ceil=800.0, frame_period=5.0, device=0):
        if not isinstance(x, torch.Tensor) or x.dtype != torch.float64 or x.dim() != 2:
            raise TypeError("x must be a float64 tensor of shape (batch, samples)")
        xs = _rows(x)
        results, record = _core._harvest_record(xs, int(fs), float(f0_floor), float(f0_ceil),
                                                float(frame_period), device=int(device))
        ctx.record, ctx.xs, ctx.device = record, xs, int(device)
        f0 = torch.from_numpy(np.stack([r[0] for r in results]))
        return f0.to(x.device)

    @staticmethod
    def backward(ctx, grad_f0):
        gs = _rows(grad_f0)
        gx = _core._harvest_adjoint(ctx.record, ctx.xs, gs, device=ctx.device)
        grad = torch.from_numpy(np.stack(gx)).to(grad_f0.device)
        return grad, None, None, None, None, None


def harvest_f0(x: torch.Tensor, fs: int, f0_floor: float = 71.0, f0_ceil: float = 800.0,
               frame_period: float = 5.0, *, device: int = 0) -> torch.Tensor:
    """Differentiable F0 (batch, frames) of float64 waveforms (batch, samples)."""
    return HarvestF0.apply(x, fs, f0_floor, f0_ceil, frame_period, device)


def frozen_f0(record, x: torch.Tensor, *, device: int = 0) -> torch.Tensor:
    """F0 of (perturbed) waveforms with a recorded forward pass's decisions
    held; record is the second result of ``_core._harvest_record``."""
    return torch.from_numpy(np.stack(_core._harvest_frozen(record, _rows(x), device=device)))
