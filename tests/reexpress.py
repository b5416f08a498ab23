"""Torch (float64, CPU) re-expression of the frozen-decision Harvest map, for
checking the hand-written adjoint with autograd.  Written from WORLD's
definitions, independent of the backend: GetMeanF0 / GetMainWindow /
GetDiffWindow / FixF0 (harvest.cpp:446-584), FixStep4 (:1000-1022),
FilteringF0 / SmoothF0Contour (:1049-1113) and the resampling of
harvest.cpp:1246-1251.  Test infrastructure only."""
import math

import numpy as np
import torch

B = (0.0078202080334971724, 0.015640416066994345)
A = (1.7347257688092754, -0.76600660094326412)


def mround(v):
    return int(v + 0.5) if v > 0 else int(v - 0.5)


def refined_f0(y, fs, frame, f0):
    """GetRefinedF0 of one slot with its discrete choices fixed by f0."""
    t = frame / 1000.0
    hw = int(1.5 * fs / f0 + 1.0)
    length = 2 * hw + 1
    n = 2 ** (2 + int(math.floor(math.log2(length))))
    basic = mround((t + (-hw) / fs) * fs + 0.001)
    m = torch.arange(length, dtype=torch.float64)
    tmp = (basic + m - 1.0) / fs - t
    wl = length / fs
    w = 0.42 + 0.5 * torch.cos(2 * math.pi * tmp / wl) + 0.08 * torch.cos(4 * math.pi * tmp / wl)
    d = torch.empty_like(w)
    d[0] = -w[1] / 2
    d[-1] = w[-2] / 2
    d[1:-1] = -(w[2:] - w[:-2]) / 2
    idx = torch.clamp(torch.arange(length) + basic - 1, 0, len(y) - 1)
    xv = y[idx]
    H = min(int(fs / 2.0 / f0), 6)
    numer = den = 0.0
    for h in range(H):
        k = mround(f0 * n / fs * (h + 1))
        ang = 2 * math.pi * ((k * np.arange(length)) % n) / n
        c = torch.tensor(np.cos(ang))
        s = torch.tensor(np.sin(ang))
        mr, mi = (xv * w * c).sum(), -(xv * w * s).sum()
        dr, di = (xv * d * c).sum(), -(xv * d * s).sum()
        p = mr * mr + mi * mi
        nu = mr * di - mi * dr
        ifr = k * fs / n + nu / p * fs / 2.0 / math.pi
        a = torch.sqrt(p)
        numer = numer + a * ifr
        den = den + a * (h + 1.0)
    return numer / (den + 1e-12)


def fix_step4(v):
    """FixStep4 on a torch contour: bridge gaps shorter than 9 frames."""
    f = len(v)
    voiced = (v.detach().numpy() > 0).astype(int)
    voiced[0] = voiced[-1] = 0
    bounds = []
    prev = 0
    for i in range(1, f):
        cur = 0 if i == f - 1 else voiced[i]
        if cur != prev:
            bounds.append(i - len(bounds) % 2)
        prev = cur
    out = v.clone()
    for s in range(len(bounds) // 2 - 1):
        a, b = bounds[2 * s + 1], bounds[2 * s + 2]
        dist = b - a - 1
        if dist >= 9:
            continue
        tmp0, tmp1 = v[a] + 1, v[b] - 1
        coef = (tmp1 - tmp0) / (dist + 1.0)
        for c, j in enumerate(range(a + 1, b), start=1):
            out = out.index_put((torch.tensor([j]),), (tmp0 + coef * c).reshape(1))
    return out


def smooth(x, lag=300):
    """SmoothF0Contour: FilteringF0 per voiced section of the padded contour."""
    mask = x.detach().numpy() > 0
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
        out = out.index_put((torch.arange(st - lag, ed - lag + 1),), torch.stack(seq[st:ed + 1]))
    return out


def frozen_f0(x, record):
    """Output F0 of one utterance at sample rate 8 kHz (no decimation) with
    the record's decisions held."""
    fs = record.fs
    assert fs == 8000 and len(record.lengths) == 1
    y = x - x.mean()
    frames = len(record.sources[0])
    src = np.asarray(record.sources[0])
    values = {}
    step3 = torch.zeros(frames, dtype=torch.float64)
    for i in np.nonzero(src >= 0)[0]:
        e = int(src[i])
        if e not in values:
            values[e] = refined_f0(y, fs, int(record.entry_frame[e]), float(record.entry_init[e]))
        step3 = step3.index_put((torch.tensor([int(i)]),), values[e].reshape(1))
    basic = smooth(fix_step4(step3))
    out_frames = record.out_frames[0]
    idx = [min(frames - 1, mround(j * record.frame_period / 1000.0 * 1000.0)) for j in range(out_frames)]
    return basic[torch.tensor(idx)]
