"""The host engine's O(F) contour stages against WORLD, bit for bit (CPU).

Counterpart of the reference's tests/cpp/test_contour.cpp:64-151: random
candidate matrices with crossable and non-crossable gaps, candidate searches
at adjacent representable values, and empty contours.  WORLD's own results
come from the golden fixtures (made from the reference's WORLD) and, where
the WORLD tap is built, live; the oracle restatement is checked alongside."""
import numpy as np
import pytest
from conftest import GOLDEN


def csr(f0, sc):
    nz = f0 != 0
    off = np.concatenate([[0], np.cumsum(nz.sum(1))]).astype(np.int32)
    return off, np.ascontiguousarray(f0[nz]), np.ascontiguousarray(sc[nz])


def engine_contour(core, oracle, f0, sc):
    off, cf, cs = csr(f0, sc)
    base = oracle.base(f0, sc, f0.shape[1])
    return core._fix_contour(off, cf, cs, base)


def test_contour_matches_world_golden(core, oracle):
    import golden.make_golden as mg
    g = np.load(GOLDEN / "world_contour.npz")
    for trial in range(40):
        f0, sc = mg.random_candidates(trial)
        c = engine_contour(core, oracle, f0, sc)
        np.testing.assert_array_equal(c, g[f"{trial}_contour"], err_msg=f"trial {trial}")
        np.testing.assert_array_equal(core._smooth(c), g[f"{trial}_smooth"], err_msg=f"trial {trial}")


def random_matrix(rng, trial, cols=8):
    pitches = np.array([80, 90, 120, 150, 180, 250, 400, 600.0])
    frames = 2 + int(rng.integers(2000))
    f0 = np.zeros((frames, cols))
    sc = np.zeros((frames, cols))
    rows = np.arange(frames)
    for col in range(cols):
        period, gap, phase = 10 + int(rng.integers(150)), int(rng.integers(8)), int(rng.integers(160))
        voiced = (trial % 4 == 0) | (((rows + phase) % period) >= gap)
        f0[:, col] = np.where(voiced, pitches[col % 4 if trial % 3 == 0 else col], 0)
        sc[:, col] = np.where(voiced, 1 + rng.integers(10, size=frames), 0)
    return f0, sc


def test_contour_matches_oracle_random(core, oracle):
    rng = np.random.default_rng(20261110)
    for trial in range(300):
        f0, sc = random_matrix(rng, trial)
        if trial % 10 == 0:
            f0[:] = 0
            sc[:] = 0
        expect = oracle.fix_contour(f0, sc, f0.shape[1])
        got = engine_contour(core, oracle, f0, sc)
        np.testing.assert_array_equal(got, expect, err_msg=f"trial {trial}")
        np.testing.assert_array_equal(core._smooth(got), oracle.smooth(got), err_msg=f"trial {trial}")


def test_contour_adjacent_values_and_ties(core, oracle):
    """ExtendF0 keeps the last candidate on equal error and applies the 18 %
    cutoff on rounded errors (harvest.cpp:636-650, 791-820)."""
    rng = np.random.default_rng(719)
    for trial in range(400):
        reference = 20.0 + rng.integers(4000000) / 1000.0
        value = reference * (0.82 if trial % 3 == 0 else 1.18 if trial % 3 == 1 else 1.001)
        frames = 40
        f0 = np.zeros((frames, 8))
        sc = np.zeros((frames, 8))
        f0[10:30, 0] = reference
        sc[10:30, 0] = 5.0
        vals = []
        for j in range(8):
            vals.append(value)
            value = np.nextafter(value, 0.0 if trial % 2 else 1e4)
        for row in (9, 30):
            f0[row, :] = rng.permutation(vals)
            sc[row, :] = rng.integers(1, 9, 8)
        expect = oracle.fix_contour(f0, sc, 8)
        np.testing.assert_array_equal(engine_contour(core, oracle, f0, sc), expect)


def test_search_score_ignores_nan_and_keeps_infinite(core, oracle):
    rng = np.random.default_rng(3)
    for trial in range(50):
        f0, sc = random_matrix(rng, trial)
        sc[rng.random(sc.shape) < 0.05] = np.nan
        sc[rng.random(sc.shape) < 0.02] = np.inf
        expect = oracle.fix_contour(f0, sc, f0.shape[1])
        got = engine_contour(core, oracle, f0, sc)
        np.testing.assert_array_equal(got, expect)


@pytest.mark.parametrize("frames", [1, 2, 3, 1000])
def test_empty_contours(core, oracle, frames):
    f0 = np.zeros((frames, 1))
    got = engine_contour(core, oracle, f0, f0.copy())
    np.testing.assert_array_equal(got, np.zeros(frames))
    np.testing.assert_array_equal(core._smooth(np.zeros(frames)), np.zeros(frames))


def test_smoothing_long_sections_settle_exactly(core, oracle):
    """Long voiced sections and long gaps exercise the settled-state skip."""
    rng = np.random.default_rng(8)
    f0 = np.zeros(20000)
    pos = 5
    while pos < len(f0) - 5:
        n = int(rng.integers(1, 3000))
        f0[pos:pos + n] = 100 + 50 * np.sin(np.arange(n) / 300.0) + rng.normal(0, 0.5, n)
        pos += n + int(rng.integers(1, 2500))
    f0[len(f0) - 5:] = 0
    np.testing.assert_array_equal(core._smooth(f0), oracle.smooth(f0))
    f0[len(f0) - 1] = 120.0  # a voiced last frame is a section of the padded contour
    np.testing.assert_array_equal(core._smooth(f0), oracle.smooth(f0))


@pytest.mark.parametrize("threads", [2, 3, 8])
def test_threaded_sections_are_bit_identical(core, oracle, threads):
    """Long contours (>= 16384 frames) extend and smooth their voiced sections
    on several threads; every output equals the single-thread result and the
    oracle's."""
    rng = np.random.default_rng(4242 + threads)
    for trial in range(3):
        f0, sc = random_matrix(rng, trial + 1)
        reps = -(-20000 // f0.shape[0])
        f0, sc = np.tile(f0, (reps, 1)), np.tile(sc, (reps, 1))
        off, cf, cs = csr(f0, sc)
        base = oracle.base(f0, sc, f0.shape[1])
        one = core._fix_contour(off, cf, cs, base)
        many = core._fix_contour(off, cf, cs, base, threads=threads)
        np.testing.assert_array_equal(many, one)
        np.testing.assert_array_equal(many, oracle.fix_contour(f0, sc, f0.shape[1]))
        np.testing.assert_array_equal(core._smooth(many, threads=threads), core._smooth(one))
        np.testing.assert_array_equal(core._smooth(many, threads=threads), oracle.smooth(one))


needs_ref = pytest.mark.skipif("not __import__('oracle.oracle').oracle.ref_available()",
                               reason="WORLD tap (oracle/_ref) not built here")


@needs_ref
def test_contour_matches_world_live(core, oracle):
    rng = np.random.default_rng(20261111)
    for trial in range(200):
        f0, sc = random_matrix(rng, trial)
        expect = oracle.ref_fix_contour(f0, sc, f0.shape[1])
        got = engine_contour(core, oracle, f0, sc)
        np.testing.assert_array_equal(got, expect, err_msg=f"trial {trial}")
        np.testing.assert_array_equal(core._smooth(got), oracle.ref_smooth(got))


def _long_contour(seed=99):
    rng = np.random.default_rng(seed)
    f0, sc = random_matrix(rng, 1)
    reps = -(-30000 // f0.shape[0])
    return np.tile(f0, (reps, 1)), np.tile(sc, (reps, 1))


def test_threaded_sections_from_concurrent_callers(core, oracle):
    """Several host threads use the section pool at once (the busy pool runs a
    caller's loop on its own thread); every result stays bit-identical."""
    from concurrent.futures import ThreadPoolExecutor
    f0, sc = _long_contour()
    off, cf, cs = csr(f0, sc)
    base = oracle.base(f0, sc, f0.shape[1])
    want = core._fix_contour(off, cf, cs, base)
    want_s = core._smooth(want)
    with ThreadPoolExecutor(max_workers=4) as pool:
        futs = [pool.submit(core._fix_contour, off, cf, cs, base, threads=4) for _ in range(6)]
        futs += [pool.submit(core._smooth, want, threads=4) for _ in range(6)]
        res = [f.result() for f in futs]
    for r in res[:6]:
        np.testing.assert_array_equal(r, want)
    for r in res[6:]:
        np.testing.assert_array_equal(r, want_s)


def _smooth_in_child(args):
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "harvest-cuda_amd"))
    from harvest_cuda import _core
    f0, threads = args
    return _core._smooth(f0, threads=threads)


def test_section_pool_survives_fork(core):
    """A forked child (e.g. a data-loader worker) starts its own pool threads
    instead of waiting for the parent's."""
    import multiprocessing as mp
    f0, sc = _long_contour(7)
    contour = np.where(f0[:, 0] > 0, f0[:, 0], 0.0)
    want = core._smooth(contour, threads=4)  # the parent's pool now has workers
    ctx = mp.get_context("fork")
    with ctx.Pool(1) as pool:
        got = pool.apply_async(_smooth_in_child, ((contour, 4),)).get(timeout=120)
    np.testing.assert_array_equal(got, want)


def test_fix_step1_threshold_ties(core, oracle):
    """FixStep1's 0.008 tests (harvest.cpp:710-722) at quotients within a few
    ulps of the threshold, zero neighbours and negative references: the
    engine's decisions equal the oracle's, bit for bit."""
    rng = np.random.default_rng(1718)
    for trial in range(300):
        frames = 12
        base = np.zeros(frames)
        b0, b1 = 100.0 + rng.random() * 300.0, 100.0 + rng.random() * 300.0
        if trial % 7 == 0:
            b0 = 0.0  # reference 2 b1 - 0
        if trial % 11 == 0:
            b0 = 3.0 * b1  # negative reference
        ref = b1 * 2 - b0
        target = ref + ref * 0.008 * (1 if trial % 2 else -1)
        for _ in range(int(rng.integers(0, 6))):
            target = np.nextafter(target, np.inf if trial % 3 else -np.inf)
        base[3:9] = [b0, b1, abs(target) if target != 0 else 1.0, 150.0, 151.0, 152.0]
        if trial % 13 == 0:
            base[4] = 0.0  # previous frame unvoiced
        f0 = base[:, None].copy()
        sc = np.where(f0 > 0, 1.0, 0.0)
        expect = oracle.fix_contour(f0, sc, 1)
        got = engine_contour(core, oracle, f0, sc)
        np.testing.assert_array_equal(got, expect, err_msg=f"trial {trial}")


def test_short_sections_at_the_edges(core, oracle):
    """FixStep1 and FixStep2 (harvest.cpp:710-762), fused into one pass in the
    engine: dense voiced runs of 1-9 frames (FixStep2 drops those spanning
    fewer than 6 frame steps), runs touching frames 1 and frames - 2, a voiced
    last frame (boundary_list counts it unvoiced) and single-frame contours,
    against the oracle's separate steps."""
    rng = np.random.default_rng(4242)
    for trial in range(400):
        frames = 1 + int(rng.integers(60)) if trial % 4 else 1 + int(rng.integers(4))
        base = np.zeros(frames)
        i = int(rng.integers(0, 3))
        while i < frames:
            run = int(rng.integers(1, 10))
            f = 120.0 + rng.random() * 5.0
            for j in range(i, min(frames, i + run)):
                base[j] = f * (1.0 + 0.001 * (j - i))  # smooth: FixStep1 keeps it
            i += run + int(rng.integers(1, 4))
        if trial % 5 == 0:
            base[-1] = 130.0
        f0 = base[:, None].copy()
        sc = np.where(f0 > 0, 1.0, 0.0)
        expect = oracle.fix_contour(f0, sc, 1)
        got = engine_contour(core, oracle, f0, sc)
        np.testing.assert_array_equal(got, expect, err_msg=f"trial {trial}")
