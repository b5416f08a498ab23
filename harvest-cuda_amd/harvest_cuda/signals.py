"""Deterministic synthetic signals for tests and benchmarks.

``config_*`` build the five benchmark workloads:

1. one 5 s signal at 16 kHz: 10-harmonic glide 100->300 Hz, 1 s silence,
   1 s white noise, 20 dB SNR noise throughout (seed 1001);
2. one 60 s signal at 48 kHz of alternating voiced/unvoiced segments
   (seed 1002);
3. 512 utterances at 24 kHz, lengths log-uniform in [0.5, 30] s, SNR
   uniform in [10, 40] dB (seed 1003);
4. 32 utterances of 4 s at 24 kHz (seed 1004);
5. 16 recordings of 30 min at 44.1 kHz with a 0.2 Hz amplitude modulation
   (seed 1005).

Free choices of the voiced/unvoiced process are fixed here: one log-F0 step
per millisecond, linear interpolation of F0 to the sample grid, harmonic k
kept only while k f(t) < fs / 2, and unvoiced noise at the noise level of the
preceding voiced segment.
"""
from __future__ import annotations

import numpy as np


def config1(seed: int = 1001) -> tuple[np.ndarray, int]:
    fs = 16000
    rng = np.random.default_rng(seed)
    n = 5 * fs
    x = np.zeros(n)
    nv = 3 * fs
    f = np.linspace(100.0, 300.0, nv)
    phi = 2 * np.pi * np.cumsum(f) / fs
    theta = rng.uniform(0, 2 * np.pi, 10)
    voiced = sum(np.sin(k * phi + theta[k - 1]) / k for k in range(1, 11))
    x[:nv] = voiced
    x[4 * fs:] = rng.normal(0, 0.05, fs)
    sigma = np.sqrt(np.mean(voiced ** 2) / 100.0)
    x += rng.normal(0, sigma, n)
    return np.ascontiguousarray(x), fs


def _voiced_unvoiced(rng: np.random.Generator, n: int, fs: int, snr_db: float,
                     harmonics: int = 20) -> np.ndarray:
    out = np.empty(n)
    pos = 0
    sigma = 0.01
    voiced = True
    while pos < n:
        if voiced:
            length = min(n - pos, int(round(rng.uniform(0.2, 2.0) * fs)))
            steps = int(np.ceil(length * 1000.0 / fs)) + 2
            start = np.log(np.exp(rng.uniform(np.log(100.0), np.log(300.0))))
            walk = np.clip(start + np.cumsum(rng.normal(0, 0.01, steps)), np.log(80.0), np.log(600.0))
            t_ms = np.arange(length) * 1000.0 / fs
            f = np.exp(np.interp(t_ms, np.arange(steps), walk))
            phi = 2 * np.pi * np.cumsum(f) / fs
            amp = rng.uniform(0.5, 1.0, harmonics) / np.arange(1, harmonics + 1)
            theta = rng.uniform(0, 2 * np.pi, harmonics)
            seg = np.zeros(length)
            # e^{i k phi} by repeated multiplication: one complex exponential
            # per sample instead of one sine per harmonic.
            base = np.exp(1j * phi)
            rot = np.ones(length, dtype=np.complex128)
            for k in range(1, harmonics + 1):
                rot *= base
                mask = k * f < fs / 2
                if not mask.any():
                    break
                seg += np.where(mask, amp[k - 1] * (rot * np.exp(1j * theta[k - 1])).imag, 0.0)
            sigma = np.sqrt(np.mean(seg ** 2) / 10 ** (snr_db / 10))
            out[pos:pos + length] = seg + rng.normal(0, sigma, length)
        else:
            length = min(n - pos, int(round(rng.uniform(0.05, 0.5) * fs)))
            out[pos:pos + length] = rng.normal(0, sigma, length)
        pos += length
        voiced = not voiced
    return out


def config2(seed: int = 1002, seconds: float = 60.0) -> tuple[np.ndarray, int]:
    fs = 48000
    rng = np.random.default_rng(seed)
    return np.ascontiguousarray(_voiced_unvoiced(rng, int(seconds * fs), fs, 25.0)), fs


def config3(seed: int = 1003, count: int = 512, lo: float = 0.5, hi: float = 30.0):
    fs = 24000
    seq = np.random.SeedSequence(seed)
    top = np.random.default_rng(seq)
    lengths = np.exp(top.uniform(np.log(lo), np.log(hi), count))
    snrs = top.uniform(10.0, 40.0, count)
    xs = []
    for child, seconds, snr in zip(seq.spawn(count), lengths, snrs):
        rng = np.random.default_rng(child)
        xs.append(np.ascontiguousarray(_voiced_unvoiced(rng, int(round(seconds * fs)), fs, snr)))
    return xs, fs


def config4(seed: int = 1004, count: int = 32):
    fs = 24000
    seq = np.random.SeedSequence(seed)
    xs = [np.ascontiguousarray(_voiced_unvoiced(np.random.default_rng(c), 4 * fs, fs, 25.0))
          for c in seq.spawn(count)]
    return xs, fs


def config4_target(f0s, seed: int = 1004, sigma: float = 0.02):
    """Per-frame targets f0 * exp(N(0, sigma^2)) for the forward F0 tracks f0s
    (the config-4 loss), from a stream independent of the waveforms'."""
    rng = np.random.default_rng(np.random.SeedSequence(seed).spawn(len(f0s) + 1)[len(f0s)])
    return [np.asarray(f) * np.exp(rng.normal(0.0, sigma, len(f))) for f in f0s]


def config5_recording(index: int, seed: int = 1005, seconds: float = 1800.0) -> tuple[np.ndarray, int]:
    fs = 44100
    child = np.random.SeedSequence(seed).spawn(index + 1)[index]
    rng = np.random.default_rng(child)
    n = int(seconds * fs)
    x = _voiced_unvoiced(rng, n, fs, 25.0)
    psi = rng.uniform(0, 2 * np.pi)
    t = np.arange(n) / fs
    x *= 0.55 + 0.45 * np.sin(2 * np.pi * 0.2 * t + psi)
    return np.ascontiguousarray(x), fs
