"""Differentiable form on the GPU (frozen decisions).

Config 4 (32 x 4 s at 24 kHz, config-2 content, seed 1004): loss = mean over
forward-voiced frames of (ln f0 - ln target)^2, target = f0 exp(N(0, 0.02^2));
Richardson-extrapolated central finite differences of the frozen-decision
forward at 96 sample indices must match the hand-written adjoint within 1e-6
relative wherever the difference resolves it.  A torch float64 CPU autograd re-expression of the same frozen map
(tests/reexpress.py) is the second check."""
import numpy as np
import pytest
import torch

import reexpress
from harvest_cuda.signals import config4, config4_target
from ref_signals import reference_signal

pytestmark = pytest.mark.gpu


def loss_terms(f0s, targets):
    """Per-frame loss terms and the loss gradient of the forward-voiced frames."""
    voiced = [t > 0 for t in targets]
    nv = sum(int(v.sum()) for v in voiced)
    terms, grads = [], []
    for f, t, v in zip(f0s, targets, voiced):
        r = np.where(v, np.log(np.where(v, f, 1.0)) - np.log(np.where(v, t, 1.0)), 0.0)
        terms.append(np.where(v, r * r, 0.0) / nv)
        grads.append(np.ascontiguousarray(np.where(v, 2.0 * r / (np.where(v, f, 1.0) * nv), 0.0)))
    return terms, grads


def loss_difference(fp, fm, target, nv):
    """Sum over voiced frames of l(fp) - l(fm), l = (ln f - ln t)^2 / nv, as
    (r+ - r-)(r+ + r-) / nv with r+ - r- = log1p((fp - fm) / fm): mean removal
    couples every sample to every frame, and differencing the rounded logs of
    thousands of nearly unchanged frames would swamp small gradients."""
    v = target > 0
    fp, fm, t = fp[v], fm[v], target[v]
    dr = np.log1p((fp - fm) / fm)
    sr = np.log(fp) + np.log(fm) - 2.0 * np.log(t)
    return float(np.sum(dr * sr)) / nv


def test_frozen_forward_reproduces_forward(gpu):
    xs, fs = config4(count=4)
    res, rec = gpu._harvest_record(xs, fs)
    frozen = gpu._harvest_frozen(rec, xs)
    for (f0, _), g in zip(res, frozen):
        assert np.array_equal(f0 > 0, g > 0)
        v = f0 > 0
        assert np.max(np.abs(g[v] - f0[v]) / f0[v]) < 1e-9
    assert rec.entries > 100


def test_config4_finite_differences_match_adjoint(gpu):
    """Per sample index, against Richardson-extrapolated central differences
    D = (4 D(h) - D(2h)) / 3 with h = 2e-4 (truncation O(h^4), about 1e-15
    relative at these amplitudes): within 1e-6 relative at every index whose
    |adj| is at least 1e-3 max|grad|, and within 1e-9 max|grad| at the
    others, over 64 seeded indices, the 16 of largest |adj| and 16 seeded
indices among those at or above 1e-3 max|grad|.

    The floor is that of FP64 differences of this loss: moving one sample
    changes the rounding of every decimated sample (four IIR passes amplify
    ulp-level differences), which moves the loss by about 5e-19, i.e. about
    2e-15 = 3e-10 max|grad| in D at this h.  A smaller h raises it as 1 / h
    (h = 1e-6 gives 6e-8 max|grad|); a larger one lets the O(h^4) term grow.
    Directional differences along random unit-variance vectors (plain central
    differences, h = 1e-6: the vector moves every sample) stand far above
    the floor and are held to 1e-6 relative alone."""
    xs, fs = config4()
    res, rec = gpu._harvest_record(xs, fs)
    f0s = [r[0] for r in res]
    targets = config4_target(f0s)
    _, g = loss_terms(f0s, targets)
    nv = sum(int((t > 0).sum()) for t in targets)
    gx = gpu._harvest_adjoint(rec, xs, g)
    rng = np.random.default_rng(np.random.SeedSequence(1004).spawn(34)[33])
    h = 2e-4
    scale = max(np.abs(a).max() for a in gx)

    def central(dirs, step):
        fp = gpu._harvest_frozen(rec, [np.ascontiguousarray(x + step * d) for x, d in zip(xs, dirs)])
        fm = gpu._harvest_frozen(rec, [np.ascontiguousarray(x - step * d) for x, d in zip(xs, dirs)])
        return sum(loss_difference(a, b, t, nv) for a, b, t in zip(fp, fm, targets)) / (2 * step)

    def fd_along(dirs):
        return (4.0 * central(dirs, h) - central(dirs, 2 * h)) / 3.0

    picks = [(int(u), int(rng.integers(len(xs[u])))) for u in rng.integers(len(xs), size=64)]
    flat = np.concatenate([np.abs(a) for a in gx])
    ends = np.cumsum([len(a) for a in gx])
    big = np.flatnonzero(flat >= 1e-3 * scale)
    for i in [*np.argsort(flat)[-16:], *rng.choice(big, size=min(16, len(big)), replace=False)]:
        u = int(np.searchsorted(ends, i, side="right"))
        picks.append((u, int(i - (ends[u - 1] if u else 0))))
    resolved, worst = 0, 0.0
    for u, k in picks:
        dirs = [np.zeros_like(x) for x in xs]
        dirs[u][k] = 1.0
        fd, adj = fd_along(dirs), float(gx[u][k])
        if abs(adj) >= 1e-3 * scale:
            resolved += 1
            worst = max(worst, abs(fd - adj) / abs(adj))
            assert abs(fd - adj) <= 1e-6 * abs(adj), (u, k, fd, adj)
        else:
            assert abs(fd - adj) <= 1e-9 * scale, (u, k, fd, adj, scale)
    assert resolved >= 32, resolved  # the relative bar covers many indices, not only the top 16
    print(f"per-index: {resolved} of {len(picks)} held to 1e-6 relative, worst {worst:.2e}")
    for _ in range(4):  # a unit-variance direction over every sample: a far smaller step
        dirs = [rng.normal(size=len(x)) for x in xs]
        fd = central(dirs, 1e-6)
        adj = sum(float(np.dot(a, d)) for a, d in zip(gx, dirs))
        assert abs(fd - adj) <= 1e-6 * abs(adj), (fd, adj)


def test_adjoint_matches_torch_reexpression(gpu):
    fs = 8000
    x = reference_signal("speechlike", seconds=1.2, fs=fs)
    res, rec = gpu._harvest_record([x], fs)
    f0 = res[0][0]
    rng = np.random.default_rng(5)
    target = f0 * np.exp(rng.normal(0, 0.02, len(f0)))
    _, g = loss_terms([f0], [target])
    gx = gpu._harvest_adjoint(rec, [x], g)[0]
    xt = torch.tensor(x, requires_grad=True)
    f = reexpress.frozen_f0(xt, rec)
    np.testing.assert_allclose(f.detach().numpy()[f0 > 0], f0[f0 > 0], rtol=1e-9)
    (f * torch.tensor(g[0])).sum().backward()
    ref = xt.grad.numpy()
    np.testing.assert_allclose(gx, ref, rtol=1e-6, atol=1e-11 * np.abs(ref).max())


def test_autograd_function_end_to_end(gpu):
    from harvest_cuda.autograd import harvest_f0
    xs, fs = config4(count=2)
    x = torch.tensor(np.stack(xs), requires_grad=True)
    f0 = harvest_f0(x, fs)
    assert f0.shape == (2, 801) and f0.dtype == torch.float64
    v = f0 > 0
    loss = torch.log(f0[v]).sum()
    loss.backward()
    assert x.grad.shape == x.shape and torch.isfinite(x.grad).all()
    res, rec = gpu._harvest_record(xs, fs)
    g = [np.ascontiguousarray(np.where(r[0] > 0, 1.0 / np.where(r[0] > 0, r[0], 1.0), 0.0)) for r in res]
    want = np.stack(gpu._harvest_adjoint(rec, xs, g))
    np.testing.assert_allclose(x.grad.numpy(), want, rtol=1e-9, atol=1e-15)
