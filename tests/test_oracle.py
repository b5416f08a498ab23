"""Pins of the oracle (the C restatement of WORLD Harvest) on the CPU.

Every oracle function is checked against at least one of: the golden vectors
written by WORLD itself (tests/golden, made by make_golden.py from the
reference's vendored WORLD), an independent re-expression (scipy / numpy FFT /
brute force), or a closed form.  Where the WORLD tap library is built (this
container) the comparison is also run live on more inputs.
"""
import hashlib

import numpy as np
import pytest
from conftest import GOLDEN

from ref_signals import reference_signal


def sha(x):
    return hashlib.sha256(np.ascontiguousarray(x, dtype=np.float64).tobytes()).hexdigest()


def assert_parity(expected, actual, rtol):
    v = expected > 0
    assert np.array_equal(v, actual > 0), f"{np.count_nonzero(v != (actual > 0))} V/UV mismatches"
    if v.any():
        rel = np.max(np.abs(actual[v] - expected[v]) / expected[v])
        assert rel <= rtol, rel


# ------------------------------------------------------------- golden WORLD
def harvest_cases():
    data = np.load(GOLDEN / "world_harvest.npz")
    keys = sorted({k.rsplit("_", 1)[0] for k in data.files})
    return [(k, data) for k in keys]


@pytest.mark.parametrize("key", [k for k, _ in harvest_cases()])
def test_oracle_harvest_matches_world_golden(oracle, key):
    data = np.load(GOLDEN / "world_harvest.npz")
    fs, kind, period = key.split("_")
    x = reference_signal(kind, seconds=0.35, fs=int(fs))
    assert sha(x) == str(data[key + "_sha"]), "signal generator drifted"
    f0, t = oracle.harvest(x, int(fs), period=float(period))
    np.testing.assert_array_equal(t, data[key + "_t"])
    assert_parity(data[key + "_f0"], f0, 1e-10)


def test_oracle_stages_match_world_golden(oracle):
    g = np.load(GOLDEN / "world_stages.npz")
    fs, seconds = int(g["fs"]), float(g["seconds"])
    x = reference_signal("speechlike", seconds=seconds, fs=fs)
    assert sha(x) == str(g["sha"])
    p = oracle.params(len(x), fs)
    y = oracle.decimate(x, fs)
    np.testing.assert_allclose(y, g["y"], rtol=0, atol=1e-13 * np.abs(g["y"]).max())
    b = oracle.boundaries(p)[int(g["band"])]
    f = oracle.filtered_band(g["y"], p, b)
    np.testing.assert_allclose(f, g["filtered"], rtol=0, atol=1e-12 * np.abs(g["filtered"]).max())
    raw = oracle.raw_candidates(g["y"], p)
    assert np.array_equal(raw > 0, g["raw"] > 0)
    np.testing.assert_allclose(raw, g["raw"], rtol=1e-10)
    cand, count = oracle.official(g["raw"], p)
    assert count == int(g["count"])
    np.testing.assert_array_equal(cand, g["cand"])
    ncand = 7 * count
    rf0, rsc = oracle.refine(g["y"], p, g["cand"], ncand)
    assert np.array_equal(rf0 > 0, g["refined_f0"] > 0)
    np.testing.assert_allclose(rf0, g["refined_f0"], rtol=1e-10)
    np.testing.assert_allclose(rsc, g["refined_score"], rtol=1e-6)
    pf0, psc = oracle.prune(g["refined_f0"], g["refined_score"], ncand)
    np.testing.assert_array_equal(pf0, g["pruned_f0"])
    np.testing.assert_array_equal(psc, g["pruned_score"])
    np.testing.assert_array_equal(oracle.fix_contour(g["pruned_f0"], g["pruned_score"], ncand),
                                  g["contour"])
    np.testing.assert_array_equal(oracle.smooth(g["contour"]), g["smooth"])


def test_oracle_contour_and_smoothing_bit_exact(oracle):
    import golden.make_golden as mg
    g = np.load(GOLDEN / "world_contour.npz")
    for trial in range(40):
        f0, sc = mg.random_candidates(trial)
        c = oracle.fix_contour(f0, sc, f0.shape[1])
        np.testing.assert_array_equal(c, g[f"{trial}_contour"])
        np.testing.assert_array_equal(oracle.smooth(c), g[f"{trial}_smooth"])


def test_oracle_edge_cases_match_world_golden(oracle):
    import golden.make_golden as mg
    g = np.load(GOLDEN / "world_edges.npz")
    for name, x, fs, lo, hi in mg.edge_cases():
        assert sha(x) == str(g[name + "_sha"]), name
        f0, t = oracle.harvest(x, fs, lo, hi)
        np.testing.assert_array_equal(t, g[name + "_t"])
        if name.startswith("silence"):
            np.testing.assert_array_equal(f0, g[name + "_f0"])
        else:
            assert_parity(g[name + "_f0"], f0, 1e-9)


# ------------------------------------------------- independent re-expressions
@pytest.mark.parametrize("fs", [16000, 24000, 44100, 96000])
def test_decimation_matches_scipy_lfilter(oracle, fs):
    from scipy.signal import lfilter
    import math
    x = reference_signal("speechlike", seconds=0.2, fs=fs)
    p = oracle.params(len(x), fs)
    r = p.ratio
    co = {2: None}
    # Coefficients of FilterForDecimate: y = b0 w + b1 w1 + b1 w2 + b0 w3,
    # w = x + a0 w1 + a1 w2 + a2 w3 (matlabfunctions.cpp:116-124).
    table = {
        2: (0.041156734567757189, -0.42599112459189636, 0.041037215479961225,
            0.16797464681802227, 0.50392394045406674),
        3: (0.95039378983237421, -0.67429146741526791, 0.15412211621346475,
            0.071221945171178636, 0.21366583551353591),
        6: (1.9715352749512141, -1.4686795689225347, 0.3893908434965701,
            0.013469181309343825, 0.040407543928031475),
        12: (2.4981398605924205, -2.1368928194784025, 0.62187513816221485,
             0.0021097275904709001, 0.0063291827714127002),
        11: (2.450743295230728, -2.06794904601978, 0.59574774438332101,
             0.0026822508007163792, 0.0080467524021491377),
    }
    a0, a1, a2, b0, b1 = table[r]
    bb, aa = [b0, b1, b1, b0], [1.0, -a0, -a1, -a2]
    lag = int(math.ceil(140.0 / r) * r)
    ext = np.concatenate([np.full(lag + 9, x[0]), x, np.full(lag + 9, x[-1])])
    once = lfilter(bb, aa, ext)
    twice = lfilter(bb, aa, once[::-1])[::-1]
    padded = len(x) + 2 * lag
    nout = (padded - 1) // r + 1
    nbeg = r - r * nout + padded
    kept = twice[np.arange(nbeg, padded + 9, r) + 8]
    expect = kept[lag // r: lag // r + p.y_length]
    got = oracle.decimate(x, fs, remove_mean=False)
    np.testing.assert_allclose(got, expect, rtol=0, atol=1e-12 * np.abs(x).max())


def test_filtered_band_matches_numpy_fft_with_overwritten_bins(oracle):
    rng = np.random.default_rng(3)
    x = reference_signal("speechlike", seconds=0.3, fs=8000)
    i = np.arange(len(x))
    x = x + 0.05 * (1 - 2 * (i % 2)) + 1e-3 * rng.standard_normal(len(x))  # excite Nyquist
    p = oracle.params(len(x), 8000)
    y = oracle.decimate(x, 8000)
    n = p.fft_size
    for band in (0, 40, p.bands - 1):
        b = oracle.boundaries(p)[band]
        h = int(p.fs_work / b * 2.0 + 0.5)
        m = np.arange(2 * h + 1)
        tmp = m / (2 * h)
        nut = (0.355768 - 0.487396 * np.cos(2 * np.pi * tmp) + 0.144232 * np.cos(4 * np.pi * tmp)
               - 0.012604 * np.cos(6 * np.pi * tmp))
        c = nut * np.cos(2 * np.pi * b * (m - h) / p.fs_work)
        S = np.fft.rfft(y, n)
        P = S * np.fft.rfft(c, n)
        # WORLD's loop leaves S[n/2] * P[n/2-1] in both of the last two bins.
        P[n // 2 - 1] = P[n // 2] = S[n // 2] * P[n // 2 - 1]
        full = np.fft.irfft(P, n)  # c2r ignores the imaginary parts of bins 0, n/2
        expect = full[h + 1: h + 1 + len(y)]
        got = oracle.filtered_band(y, p, b)
        np.testing.assert_allclose(got, expect, rtol=0, atol=1e-12 * np.abs(expect).max())


def _crossings(v, fs):
    """ZeroCrossingEngine (harvest.cpp:162-197) written with numpy."""
    i = np.nonzero((v[:-1] > 0) & (v[1:] <= 0))[0]
    e = (i + 1).astype(np.float64)
    fine = e - v[i] / (v[i + 1] - v[i])
    if len(fine) < 2:
        return np.zeros(0), np.zeros(0)
    return (fine[:-1] + fine[1:]) / 2.0 / fs, fs / np.diff(fine)


def _interp1(x, y, xi):
    """interp1 with histc's clamped bracket and linear extrapolation
    (matlabfunctions.cpp:136-176); numpy.interp would clamp instead."""
    k = np.clip(np.searchsorted(x, xi, side="right"), 1, len(x) - 1)
    s = (xi - x[k - 1]) / (x[k] - x[k - 1])
    return y[k - 1] + s * (y[k] - y[k - 1])


@pytest.mark.parametrize("band", [10, 60, 120])
def test_band_candidates_match_brute_force(oracle, band):
    x = reference_signal("speechlike", seconds=0.4, fs=16000)
    p = oracle.params(len(x), 16000)
    y = oracle.decimate(x, 16000)
    b = oracle.boundaries(p)[band]
    f = oracle.filtered_band(y, p, b)
    d = f[1:] - f[:-1]
    events = [_crossings(f, p.fs_work), _crossings(-f, p.fs_work),
              _crossings(d, p.fs_work), _crossings(-d, p.fs_work)]
    t = np.arange(p.frames) * 1 / 1000.0
    if any(len(e[0]) < 3 for e in events):
        expect = np.zeros(p.frames)
    else:
        v = [_interp1(loc, val, t) for loc, val in events]
        expect = (v[0] + v[1] + v[2] + v[3]) / 4.0
        bad = (expect > b * 1.1) | (expect < b * 0.9) | (expect > 800.0) | (expect < 71.0)
        expect[bad] = 0.0
    got = oracle.band_candidates(y, p, b)
    assert np.array_equal(got > 0, expect > 0)
    np.testing.assert_allclose(got, expect, rtol=1e-12)


def test_official_and_overlap_brute_force(oracle):
    rng = np.random.default_rng(11)
    p = oracle.params(16000, 16000)
    raw = np.where(rng.random((p.bands, p.frames)) < 0.7, rng.uniform(80, 400, (p.bands, p.frames)), 0)
    cand, count = oracle.official(raw, p)
    expect = np.zeros((p.frames, p.slots))
    counts = []
    for i in range(p.frames):
        vuv = (raw[:, i] > 0).astype(int)
        vuv[0] = vuv[-1] = 0
        runs, st = [], None
        for j in range(1, p.bands):
            if vuv[j] - vuv[j - 1] == 1:
                st = j
            if vuv[j] - vuv[j - 1] == -1:
                if j - st >= 10:
                    s = 0.0
                    for q in range(st, j):
                        s += raw[q, i]
                    runs.append(s / (j - st))
        expect[i, :len(runs)] = runs
        counts.append(len(runs))
    assert count == max(counts)
    for g in range(1, 4):
        for j in range(count):
            expect[g:, j + count * g] = expect[:-g, j]
            expect[:-g, j + count * (g + 3)] = expect[g:, j]
    np.testing.assert_array_equal(cand, expect)


def test_refinement_matches_numpy_full_fft(oracle):
    """GetMeanF0 (harvest.cpp:541-584) with full numpy FFTs instead of the
    oracle's DFT at the harmonic bins."""
    rng = np.random.default_rng(5)
    fs = 8000.0
    y = reference_signal("speechlike", seconds=0.5, fs=8000)
    for _ in range(25):
        t = rng.integers(50, 450) / 1000.0
        f0 = rng.uniform(75, 780)
        h = int(1.5 * fs / f0 + 1.0)
        wl = (2.0 * h + 1.0) / fs
        n = int(2.0 ** (2 + int(np.log(h * 2.0 + 1.0) / 0.69314718055994529)))
        basic = int((t + (-h) / fs) * fs + 0.001 + 0.5)
        m = np.arange(2 * h + 1)
        tmp = (basic + m - 1.0) / fs - t
        w = 0.42 + 0.5 * np.cos(2.0 * np.pi * tmp / wl) + 0.08 * np.cos(4.0 * np.pi * tmp / wl)
        dw = np.empty_like(w)
        dw[0], dw[-1] = -w[1] / 2.0, w[-2] / 2.0
        dw[1:-1] = -(w[2:] - w[:-2]) / 2.0
        xs = y[np.clip(basic + m - 1, 0, len(y) - 1)]
        M, D = np.fft.rfft(xs * w, n), np.fft.rfft(xs * dw, n)
        H = min(int(fs / 2.0 / f0), 6)
        amp, ifr = [], []
        for k in range(H):
            idx = int(f0 * n / fs * (k + 1) + 0.5)
            pw = M[idx].real ** 2 + M[idx].imag ** 2
            nu = M[idx].real * D[idx].imag - M[idx].imag * D[idx].real
            ifr.append(0.0 if pw == 0 else idx * fs / n + nu / pw * fs / 2.0 / np.pi)
            amp.append(np.sqrt(pw))
        amp, ifr = np.array(amp), np.array(ifr)
        refined = np.sum(amp * ifr) / (np.sum(amp * np.arange(1, H + 1)) + 1e-12)
        score = 1.0 / (np.sum(np.abs((ifr / np.arange(1, H + 1) - f0) / f0)) / H + 1e-12)
        if refined < 71 or refined > 800 or score < 2.5:
            refined = score = 0.0
        r, s = oracle.refine_one(y, fs, t, f0)
        assert (r > 0) == (refined > 0)
        if r > 0:
            assert abs(r - refined) <= 1e-10 * refined
            assert abs(s - score) <= 1e-6 * score


def test_refinement_harmonic_tone_closed_form(oracle):
    """The instantaneous frequency of harmonic k of a stationary harmonic tone
    is k f, so FixF0's amplitude-weighted estimate sum a_k IF_k / sum a_k k is
    f (up to the leakage of the window's far sidelobes)."""
    fs = 8000.0
    for f in (90.0, 150.0, 333.3, 600.0):
        t = np.arange(8000) / fs
        H = min(int(fs / 2.0 / (f * 1.01)), 6)
        y = sum(np.sin(2 * np.pi * k * f * t + 0.3 * k) / k for k in range(1, H + 1))
        r, s = oracle.refine_one(y, fs, 0.5, f * 1.01)
        assert abs(r - f) / f < 5e-4, (f, r)
        assert s > 2.5


def test_prune_and_base_brute_force(oracle):
    rng = np.random.default_rng(2)
    F, C = 60, 14
    f0 = np.where(rng.random((F, C)) < 0.5, rng.uniform(100, 120, (F, C)), 0.0)
    sc = np.where(f0 > 0, rng.uniform(3, 20, (F, C)), 0.0)
    pf, ps = oracle.prune(f0, sc, C)
    ef, es = f0.copy(), sc.copy()
    for i in range(1, F - 1):
        for j in range(C):
            r = f0[i, j]
            if r == 0:
                continue
            e1 = min([1.0] + [abs(r - c) / r for c in f0[i + 1]])
            e2 = min([1.0] + [abs(r - c) / r for c in f0[i - 1]])
            if min(e1, e2) > 0.05:
                ef[i, j] = es[i, j] = 0.0
    np.testing.assert_array_equal(pf, ef)
    np.testing.assert_array_equal(ps, es)
    base = oracle.base(pf, ps, C)
    for i in range(F):
        best, b = 0.0, 0.0
        for j in range(C):
            if ps[i, j] > best:
                best, b = ps[i, j], pf[i, j]
        assert base[i] == b


def test_params_closed_forms(oracle):
    p = oracle.params(80000, 16000, 71.0, 800.0, 5.0)
    assert (p.bands, p.ratio, p.y_length, p.fft_size, p.frames, p.out_frames, p.slots) == \
        (152, 2, 40000, 65536, 5001, 1001, 105)
    p = oracle.params(2880000, 48000)
    assert (p.y_length, p.fft_size, p.frames) == (480000, 524288, 60001)
    p = oracle.params(79380000, 44100)
    assert (p.ratio, p.fs_work, p.y_length, p.fft_size) == (6, 7350.0, 13230000, 16777216)


# ------------------------------------------------------- live against WORLD
needs_ref = pytest.mark.skipif("not __import__('oracle.oracle').oracle.ref_available()",
                               reason="WORLD tap (oracle/_ref) not built here")


@needs_ref
@pytest.mark.parametrize("fs", [8000, 16000, 24000, 44100, 48000])
@pytest.mark.parametrize("period", [1.0, 2.5, 5.0, 10.0])
def test_oracle_matches_world_live(oracle, fs, period):
    x = reference_signal("speechlike", seconds=0.35, fs=fs)
    a, ta = oracle.harvest(x, fs, period=period)
    b, tb = oracle.ref_harvest(x, fs, period=period)
    np.testing.assert_array_equal(ta, tb)
    assert_parity(b, a, 1e-10)
