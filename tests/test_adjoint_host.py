"""Host transposes of the differentiable form (CPU).

Smoothing (SmoothF0Contour, harvest.cpp:1079-1113) is linear for a fixed
voiced pattern and FixStep4 (harvest.cpp:1000-1022) is affine, so their
hand-written transposes are pinned by the dot-product identity
<A x, g> = <x, A^T g>, and the smoothing transpose also by autograd through a
torch re-expression of WORLD's FilteringF0 (harvest.cpp:1049-1074)."""
import numpy as np
import pytest
import torch

B = (0.0078202080334971724, 0.015640416066994345)
A = (1.7347257688092754, -0.76600660094326412)


def voiced_pattern(rng, frames, sections):
    mask = np.zeros(frames, bool)
    for _ in range(sections):
        st = int(rng.integers(1, frames - 2))
        mask[st:st + int(rng.integers(1, frames // 4))] = True
    mask[0] = mask[-1] = False
    return mask


def torch_smooth(x, mask, lag=300):
    """FilteringF0 per voiced section of the padded contour, in torch."""
    n = len(mask) + 2 * lag
    pm = np.zeros(n, bool)
    pm[lag:lag + len(mask)] = mask
    xp = torch.cat([torch.zeros(lag, dtype=x.dtype), x, torch.zeros(lag, dtype=x.dtype)])
    out = torch.zeros_like(x)
    i = 1
    while i < n - 1:
        if not pm[i]:
            i += 1
            continue
        st = i
        while i + 1 < n - 1 and pm[i + 1]:
            i += 1
        ed = i
        i += 1
        seq = [xp[st]] * st + [xp[k] for k in range(st, ed + 1)] + [xp[ed]] * (n - 1 - ed)
        for _ in range(2):
            w0 = w1 = torch.zeros((), dtype=x.dtype)
            res = []
            for v in seq:
                wt = v + A[0] * w0 + A[1] * w1
                res.append(B[0] * wt + B[1] * w0 + B[0] * w1)
                w1, w0 = w0, wt
            seq = res[::-1]
        out = out.index_put((torch.arange(st - lag, ed - lag + 1),),
                            torch.stack(seq[st:ed + 1]))
    return out


def test_smoothing_transpose_dot_product(core):
    rng = np.random.default_rng(11)
    for trial in range(5):
        frames = 3000
        mask = voiced_pattern(rng, frames, 6)
        x1 = np.where(mask, rng.uniform(80, 400, frames), 0.0)
        x2 = np.where(mask, rng.uniform(80, 400, frames), 0.0)
        g = rng.normal(size=frames)
        lhs = np.dot(core._smooth(x1) - core._smooth(x2), g)
        rhs = np.dot(x1 - x2, core._smooth_adjoint(x1, g))
        assert abs(lhs - rhs) <= 1e-10 * np.abs(x1 - x2).sum() * np.abs(g).max(), trial


def test_smoothing_transpose_matches_torch_autograd(core):
    rng = np.random.default_rng(12)
    frames = 400
    mask = voiced_pattern(rng, frames, 3)
    mask[50] = True
    mask[49] = mask[51] = False  # a one-frame section
    x = np.where(mask, rng.uniform(80, 400, frames), 0.0)
    g = rng.normal(size=frames)
    xt = torch.tensor(x, requires_grad=True)
    out = torch_smooth(xt, mask)
    np.testing.assert_allclose(out.detach().numpy(), core._smooth(x), rtol=1e-13, atol=1e-10)
    (out * torch.tensor(g)).sum().backward()
    grad = xt.grad.numpy() * mask
    np.testing.assert_allclose(core._smooth_adjoint(x, g), grad, rtol=1e-10, atol=1e-12)


def test_fix_step4_transpose_dot_product(core):
    rng = np.random.default_rng(13)
    for trial in range(20):
        frames = 500
        mask = np.zeros(frames, bool)
        pos = 3
        while pos < frames - 20:
            run = int(rng.integers(6, 40))
            mask[pos:pos + run] = True
            pos += run + int(rng.integers(1, 14))  # gaps of 1..13: some bridged, some not
        x1 = np.where(mask, rng.uniform(80, 400, frames), 0.0)
        x2 = np.where(mask, rng.uniform(80, 400, frames), 0.0)
        g = rng.normal(size=frames)
        lhs = np.dot(core._fix_step4(x1) - core._fix_step4(x2), g)
        rhs = np.dot(x1 - x2, core._fix_step4_adjoint(x1, g))
        assert abs(lhs - rhs) <= 1e-9 * (1 + abs(lhs)), trial
        assert np.count_nonzero(core._fix_step4(x1)) > np.count_nonzero(x1)  # gaps were bridged
