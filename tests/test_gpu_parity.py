"""GPU parity: the gfx950 path against the oracle and against WORLD's golden
vectors (needs an MI355X).

The bar: identical voiced/unvoiced decisions and a relative F0 error of at
most 1e-6 on voiced frames (the observed error is ~1e-14).  Inputs follow the
reference's tests/test_compatibility.py:25-371 (all sample rates and signal
kinds, parameters, read-only input, short silences, DC inputs, every
decimation ratio, the dispatch boundary, Nyquist-bin excitation, refinement
window sizes, separate voiced sections, concurrency and cache clearing)."""
import hashlib
import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
from conftest import GOLDEN

from harvest_cuda.signals import config1, config2
from ref_signals import reference_signal

pytestmark = pytest.mark.gpu
RTOL = 1e-6


def sha(x):
    return hashlib.sha256(np.ascontiguousarray(x, dtype=np.float64).tobytes()).hexdigest()


def assert_parity(expected, actual, rtol=RTOL):
    v = expected > 0
    assert expected.shape == actual.shape
    assert np.array_equal(v, actual > 0), f"{np.count_nonzero(v != (actual > 0))} V/UV mismatches"
    if v.any():
        rel = np.max(np.abs(actual[v] - expected[v]) / expected[v])
        assert rel <= rtol, rel


@pytest.mark.parametrize("fs", [8000, 16000, 24000, 44100, 48000])
@pytest.mark.parametrize("kind", ["speechlike", "low", "high", "noise", "silence"])
def test_matches_oracle(gpu, oracle, fs, kind):
    x = reference_signal(kind, seconds=0.7, fs=fs)
    f0, t = gpu.harvest(x, fs)
    ef0, et = oracle.harvest(x, fs)
    assert f0.dtype == np.float64
    np.testing.assert_array_equal(t, et)
    assert_parity(ef0, f0)
    if kind in ("low", "high", "speechlike"):
        assert np.count_nonzero(ef0) > 10


def golden_keys():
    data = np.load(GOLDEN / "world_harvest.npz")
    return sorted({k.rsplit("_", 1)[0] for k in data.files})


@pytest.mark.parametrize("key", golden_keys())
def test_matches_world_golden(gpu, key):
    data = np.load(GOLDEN / "world_harvest.npz")
    fs, kind, period = key.split("_")
    x = reference_signal(kind, seconds=0.35, fs=int(fs))
    assert sha(x) == str(data[key + "_sha"])
    f0, t = gpu.harvest(x, int(fs), frame_period=float(period))
    np.testing.assert_array_equal(t, data[key + "_t"])
    assert_parity(data[key + "_f0"], f0)


def test_edge_cases_match_world_golden(gpu):
    """Short silences (exact), every decimation ratio with a partial final
    group, refinement windows from 8 Hz floors to 2.4 kHz tones at 8 kHz and
    192 kHz input, and excitation of the two overwritten Nyquist bins."""
    import golden.make_golden as mg
    g = np.load(GOLDEN / "world_edges.npz")
    for name, x, fs, lo, hi in mg.edge_cases():
        assert sha(x) == str(g[name + "_sha"]), name
        f0, t = gpu.harvest(x, fs, lo, hi)
        np.testing.assert_array_equal(t, g[name + "_t"], err_msg=name)
        if name.startswith("silence"):
            np.testing.assert_array_equal(f0, g[name + "_f0"], err_msg=name)
        else:
            assert_parity(g[name + "_f0"], f0)


@pytest.mark.parametrize("period,floor,ceil", [(1.0, 50, 900), (2.5, 100, 400), (7.3, 65, 1000)])
def test_parameters_and_readonly_input(gpu, oracle, period, floor, ceil):
    x = reference_signal(seconds=1.1, fs=44100)
    x.flags.writeable = False
    f0, t = gpu.harvest(x, 44100, floor, ceil, period)
    ef0, et = oracle.harvest(x, 44100, floor, ceil, period)
    np.testing.assert_array_equal(t, et)
    assert_parity(ef0, f0)


@pytest.mark.parametrize("fs", [72000, 80000, 192000])
@pytest.mark.parametrize("constant", [True, False])
def test_dc_input(gpu, fs, constant):
    x = np.ones(round(2.2 * fs)) if constant else 1000 + 0.01 * reference_signal(seconds=2.2, fs=fs)
    f0, t = gpu.harvest(x, fs)
    assert f0.shape == t.shape
    assert np.all(np.isfinite(f0)) and np.all(f0 >= 0)


@pytest.mark.parametrize("length", [32767, 32768, 32769])
def test_length_boundaries(gpu, oracle, length):
    x = reference_signal(seconds=length / 24000, fs=24000)
    f0, t = gpu.harvest(x, 24000)
    ef0, et = oracle.harvest(x, 24000)
    np.testing.assert_array_equal(t, et)
    assert_parity(ef0, f0)


def test_separate_voiced_sections(gpu, oracle):
    fs = 16000
    t = np.arange(fs * 8) / fs
    x = sum(np.sin(2 * np.pi * 137 * h * t) / h for h in range(1, 6))
    x[(t % 0.8 > 0.5) & (t < 7.5)] = 0
    f0, _ = gpu.harvest(x, fs)
    ef0, _ = oracle.harvest(x, fs)
    assert np.count_nonzero(ef0) > 100 and np.count_nonzero(ef0 == 0) > 100
    assert_parity(ef0, f0)


@pytest.mark.parametrize("fs", [8000, 16000, 24000, 44100, 48000, 96000, 192000])
def test_stage_decimation(gpu, oracle, fs):
    x = reference_signal(seconds=0.9, fs=fs)
    got = gpu._decimate(x, fs)
    want = oracle.decimate(x, fs, remove_mean=False)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(x).max())


@pytest.mark.parametrize("fs,kind", [(16000, "speechlike"), (24000, "low"), (48000, "high")])
def test_stage_candidates(gpu, oracle, fs, kind):
    """Pruned refined candidates (CSR) and best F0 per 1 ms frame against the
    oracle's RemoveUnreliableCandidates / SearchF0Base outputs."""
    x = reference_signal(kind, seconds=0.7, fs=fs)
    off, f0, sc, best = gpu._candidates([x], fs)
    st = oracle.stages(x, fs)
    ref = st["f0"][:, :st["ncand"]]
    refs = st["score"][:, :st["ncand"]]
    nz = ref != 0
    np.testing.assert_array_equal(np.diff(off), nz.sum(1))
    np.testing.assert_allclose(f0, ref[nz], rtol=1e-9)
    np.testing.assert_allclose(sc, refs[nz], rtol=1e-6)
    assert_parity(st["best"], best)


def test_batched_equals_plain_bitwise(gpu):
    rng = np.random.default_rng(7)
    xs = [reference_signal(k, seconds=s, fs=24000) * a
          for k, s, a in [("speechlike", 0.3, 1.0), ("low", 1.7, 0.5), ("high", 0.05, 2.0),
                          ("noise", 0.9, 1.0), ("silence", 0.2, 1.0), ("speechlike", 3.1, 0.1)]]
    xs.append(np.ascontiguousarray(rng.normal(0, 0.1, 24000 * 2)))
    batched = gpu.harvest_batch(xs, 24000)
    for x, (bf0, bt) in zip(xs, batched):
        f0, t = gpu.harvest(x, 24000)
        np.testing.assert_array_equal(bf0, f0)
        np.testing.assert_array_equal(bt, t)


def test_device_resident_entry_matches(gpu):
    import torch
    xs = [reference_signal("speechlike", seconds=s, fs=16000) for s in (0.4, 1.3)]
    flat = torch.from_numpy(np.concatenate(xs)).to("cuda:0")
    torch.cuda.synchronize()
    out = gpu.harvest_device(flat.data_ptr(), [len(x) for x in xs], 16000)
    for x, (f0, t) in zip(xs, out):
        ef0, et = gpu.harvest(x, 16000)
        np.testing.assert_array_equal(f0, ef0)
        np.testing.assert_array_equal(t, et)


def test_deterministic_and_power_of_two_scale_invariant(gpu):
    """Repeated calls are bitwise equal.  Scaling by a power of two is exact
    through every linear stage; only FixF0's additive 1e-12 guards
    (harvest.cpp:531-532) see the scale, so the track moves by ~1e-13 and
    every voicing decision stays put."""
    x, fs = config1()
    a, _ = gpu.harvest(x, fs)
    b, _ = gpu.harvest(x, fs)
    np.testing.assert_array_equal(a, b)
    for scale in (4.0, 1 / 1024.0):
        c, _ = gpu.harvest(x * scale, fs)
        assert_parity(a, c, 1e-9)


def test_config1_matches_oracle(gpu, oracle):
    x, fs = config1()
    f0, t = gpu.harvest(x, fs)
    ef0, et = oracle.harvest(x, fs)
    np.testing.assert_array_equal(t, et)
    assert_parity(ef0, f0)
    assert np.count_nonzero(ef0) > 300


def test_config2_matches_oracle(gpu, oracle):
    x, fs = config2()
    f0, t = gpu.harvest(x, fs)
    ef0, et = oracle.harvest(x, fs)
    np.testing.assert_array_equal(t, et)
    assert_parity(ef0, f0)
    assert np.count_nonzero(ef0) > 5000


def test_profile_keys(gpu):
    x = reference_signal("speechlike", seconds=0.35, fs=16000)
    actual = gpu.harvest(x, 16000)
    f0, t, timings = gpu._profile(x, 16000)
    np.testing.assert_array_equal(f0, actual[0])
    np.testing.assert_array_equal(t, actual[1])
    for key in ("preprocess", "filter_events", "refine", "contour", "smooth", "resample"):
        assert key in timings and np.isfinite(timings[key]) and timings[key] >= 0
    assert "filter" not in timings and "events" not in timings


def test_explicit_device_and_out_of_range(gpu):
    x = reference_signal(seconds=0.35, fs=16000)
    expected = gpu.harvest(x, 16000)
    assert gpu.cuda_available(device=0)
    assert not gpu.cuda_available(device=2 ** 31 - 1)
    with pytest.raises(ValueError, match="out of range"):
        gpu.harvest(x, 16000, device=2 ** 31 - 1)
    for got in (gpu.harvest(x, 16000, device=0), gpu._profile(x, 16000, device=0)[:2]):
        for a, b in zip(got, expected):
            np.testing.assert_array_equal(a, b)


def test_concurrent_calls_and_clear_cache(gpu):
    from harvest_cuda import clear_cache
    inputs = [(reference_signal(seconds=0.3, fs=16000), 16000),
              (reference_signal("high", seconds=0.5, fs=24000), 24000)]
    expected = [gpu.harvest(x, fs) for x, fs in inputs]
    with ThreadPoolExecutor(max_workers=3) as pool:
        futures = [pool.submit(gpu.harvest, x, fs) for _ in range(3) for x, fs in inputs]
        pool.submit(clear_cache).result()
        results = [f.result() for f in futures]
    clear_cache()
    results.append(gpu.harvest(*inputs[0]))
    for k, r in enumerate(results[:-1]):
        for a, b in zip(r, expected[k % 2]):
            np.testing.assert_array_equal(a, b)
    for a, b in zip(results[-1], expected[0]):
        np.testing.assert_array_equal(a, b)


def test_clear_cache_while_a_call_runs(gpu):
    from harvest_cuda import clear_cache
    x = reference_signal(seconds=0.7, fs=24000)
    expected = gpu.harvest(x, 24000)
    ready = threading.Barrier(2)

    def worker():
        before = gpu.harvest(x, 24000)
        ready.wait(timeout=60)
        ready.wait(timeout=60)
        return before, gpu.harvest(x, 24000)

    with ThreadPoolExecutor(max_workers=1) as pool:
        fut = pool.submit(worker)
        ready.wait(timeout=60)
        clear_cache()
        clear_cache()
        ready.wait(timeout=60)
        got = [*fut.result(), gpu.harvest(x, 24000)]
    clear_cache()
    for r in got:
        for a, b in zip(r, expected):
            np.testing.assert_array_equal(a, b)


def test_c_abi_async_entry_points_match_sync(gpu):
    """hv_analyze_batch_async returns before the candidate download completes;
    after hv_wait_candidates the zero-copy view holds exactly what the
    synchronous entry point produces (called through the C ABI)."""
    import ctypes
    from pathlib import Path
    lib = ctypes.CDLL(str(Path(gpu.__file__).parent / "libharvest_hip.so"))
    P, D, I64 = ctypes.c_void_p, ctypes.c_double, ctypes.c_int64
    sig = [P, ctypes.c_int, P, P, ctypes.c_int, D, D, P, P, P, P]
    for name in ("hv_analyze_batch", "hv_analyze_batch_async"):
        getattr(lib, name).argtypes = sig
    lib.hv_wait_candidates.argtypes = [P, ctypes.c_int]
    lib.hv_fetch_candidates.argtypes = [P, P, P]
    lib.hv_result_view.argtypes = [P, P, P, P, P, P]
    lib.hv_workspace_destroy.argtypes = [P]
    fs = 16000
    xs = [reference_signal("speechlike", seconds=s, fs=fs) for s in (0.6, 1.1)]
    frames = sum(int(1000.0 * len(x) / fs) + 1 for x in xs)
    ptrs = (ctypes.c_void_p * 2)(*[x.ctypes.data for x in xs])
    lens = (I64 * 2)(*[len(x) for x in xs])
    ws = ctypes.c_void_p()
    assert lib.hv_workspace_create(0, ctypes.byref(ws)) == 0
    try:
        def analyze(fn):
            off = np.zeros(frames + 1, np.int32)
            best = np.zeros(frames)
            nnz = I64()
            assert fn(ws, 2, ptrs, lens, fs, 71.0, 800.0, off.ctypes.data, best.ctypes.data,
                      ctypes.byref(nnz), None) == 0
            return off, best, nnz.value

        def fetch(n):
            f0, sc = np.zeros(max(n, 1)), np.zeros(max(n, 1))
            assert lib.hv_fetch_candidates(ws, f0.ctypes.data, sc.ctypes.data) == 0
            return f0[:n], sc[:n]

        off1, best1, n1 = analyze(lib.hv_analyze_batch)
        f1, s1 = fetch(n1)
        assert n1 > 0
        # Deferred download, waited for explicitly, read in place.
        off2, best2, n2 = analyze(lib.hv_analyze_batch_async)
        assert lib.hv_wait_candidates(ws, 0) == 0
        assert lib.hv_wait_candidates(ws, 1) == 0
        vo, vb, vf, vs, vn = P(), P(), P(), P(), I64()
        assert lib.hv_result_view(ws, ctypes.byref(vo), ctypes.byref(vb), ctypes.byref(vf),
                                  ctypes.byref(vs), ctypes.byref(vn)) == 0
        assert vn.value == n2 == n1
        f2 = np.ctypeslib.as_array(ctypes.cast(vf, ctypes.POINTER(D)), (n2,)).copy()
        s2 = np.ctypeslib.as_array(ctypes.cast(vs, ctypes.POINTER(D)), (n2,)).copy()
        for a, b in ((off1, off2), (best1, best2), (f1, f2), (s1, s2)):
            np.testing.assert_array_equal(a, b)
        # hv_fetch_candidates waits by itself after an async call.
        analyze(lib.hv_analyze_batch_async)
        f3, s3 = fetch(n1)
        np.testing.assert_array_equal(f3, f1)
        np.testing.assert_array_equal(s3, s1)
        assert lib.hv_wait_candidates(ws, 2) == 1  # HV_EINVAL
        # Two-step form: submit returns at once, analyses wait for collect.
        lib.hv_analyze_batch_submit.argtypes = [P, ctypes.c_int, P, P, ctypes.c_int, D, D,
                                                ctypes.c_int]
        lib.hv_analyze_batch_collect.argtypes = [P, P, P, P, P]
        assert lib.hv_analyze_batch_submit(ws, 2, ptrs, lens, fs, 71.0, 800.0, 0) == 0
        nnz = I64()
        assert lib.hv_analyze_batch(ws, 2, ptrs, lens, fs, 71.0, 800.0, None, None,
                                    ctypes.byref(nnz), None) == 1  # collect first
        off4, best4 = np.zeros(frames + 1, np.int32), np.zeros(frames)
        assert lib.hv_analyze_batch_collect(ws, off4.ctypes.data, best4.ctypes.data,
                                            ctypes.byref(nnz), None) == 0
        f4, s4 = fetch(nnz.value)
        for a, b in ((off1, off4), (best1, best4), (f1, f4), (s1, s4)):
            np.testing.assert_array_equal(a, b)
        assert lib.hv_analyze_batch_collect(ws, None, None, ctypes.byref(nnz), None) == 1
    finally:
        lib.hv_workspace_destroy(ws)


def test_pipelined_batches_equal_plain_bitwise(gpu):
    """A batch of >= 4 utterances and >= 200 000 frames is analysed in groups
    alternating between two workspaces (the host contours of one group
    overlap the device work of the next); every result still equals the
    plain call's, bit for bit, for host and device-resident input."""
    import torch
    fs = 16000
    kinds = ["speechlike", "low", "noise", "speechlike", "high", "speechlike", "low"]
    secs = [41.0, 33.5, 27.0, 45.25, 30.0, 38.0, 12.0]
    xs = [reference_signal(k, seconds=s, fs=fs) for k, s in zip(kinds, secs)]
    frames = [int(1000.0 * len(x) / fs) + 1 for x in xs]
    assert sum(frames) >= 200000
    plain = [gpu.harvest(x, fs) for x in xs]
    for _ in range(2):  # the second call reuses both cached workspaces
        batched = gpu.harvest_batch(xs, fs)
        for (bf0, bt), (f0, t) in zip(batched, plain):
            np.testing.assert_array_equal(bf0, f0)
            np.testing.assert_array_equal(bt, t)
    flat = torch.from_numpy(np.concatenate(xs)).to("cuda:0")
    torch.cuda.synchronize()
    out = gpu.harvest_device(flat.data_ptr(), [len(x) for x in xs], fs)
    for (df0, dt), (f0, t) in zip(out, plain):
        np.testing.assert_array_equal(df0, f0)
        np.testing.assert_array_equal(dt, t)


def test_nonfinite_samples_rejected_by_staging(gpu):
    """With a GPU the sample check runs in the backend's staging pass (host
    input goes to pinned memory in 4 MiB slices): a NaN or inf in a later
    slice, at either end, or in a later utterance of a batch raises
    ValueError, and the next call on the same workspace is unaffected."""
    fs = 16000
    x = reference_signal("speechlike", seconds=45.0, fs=fs)  # 720 000 samples: two slices
    want_f0, want_t = gpu.harvest(x, fs)
    for bad, at in ((np.nan, 600_000), (np.inf, 0), (-np.inf, len(x) - 1)):
        y = x.copy()
        y[at] = bad
        with pytest.raises(ValueError, match="finite"):
            gpu.harvest(y, fs)
        with pytest.raises(ValueError, match="finite"):
            gpu.harvest_batch([np.ascontiguousarray(x[:50_000]), y], fs)
        f0, t = gpu.harvest(x, fs)
        np.testing.assert_array_equal(f0, want_f0)
        np.testing.assert_array_equal(t, want_t)


@pytest.mark.parametrize("floor,fs", [(30.0, 16000), (14.0, 16000), (71.0, 8000), (71.0, 96000)])
def test_filterbank_fft_and_direct_bands_match_oracle(gpu, oracle, floor, fs):
    """The long bands run by overlap-save FFT (h >= 64, longest h <= 1000),
    the rest by the direct FIR; a 14 Hz floor makes the longest band exceed
    the FFT block, so every band takes the direct FIR.  Candidates and best
    F0 against the oracle's stages either way."""
    x = reference_signal("speechlike", seconds=0.9, fs=fs)
    off, f0, sc, best = gpu._candidates([x], fs, floor, 800.0)
    st = oracle.stages(x, fs, floor, 800.0)
    ref = st["f0"][:, :st["ncand"]]
    refs = st["score"][:, :st["ncand"]]
    nz = ref != 0
    np.testing.assert_array_equal(np.diff(off), nz.sum(1))
    np.testing.assert_allclose(f0, ref[nz], rtol=1e-9)
    np.testing.assert_allclose(sc, refs[nz], rtol=1e-6)
    assert_parity(st["best"], best)
