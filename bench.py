#!/usr/bin/env python
"""Throughput benchmark of the MI355X Harvest path.

Primary line: the batched workload (512 float64 utterances at 24 kHz,
lengths log-uniform over 0.5-30 s, seed 1003) analysed by one batched call,
NumPy in and NumPy out (the reference benchmark's scope,
benchmarks/benchmark.py:73: host input, upload, device analysis, host
contour, output arrays).  One step = one harvest_batch call over the whole
workload.  `value` = audio-seconds per second over the K timed steps.
Reported beside it:
  * plain: the same utterances through one harvest() call each (the plain
    per-call GPU path, same scope), median of --plain-repeats loops, and the
    batched/plain ratio;
  * first_call: the cold first batched call (allocations included);
  * roofline of the dominant kernel from hipEvents recorded around every
    kernel launch on the backends' streams during the timed steps;
  * cpu_baseline: WORLD Harvest (built from the reference into oracle/_ref,
    or the oracle's C restatement where that is absent) threaded over the
    512 utterances, longest first, on the host's cores; its outputs are also
    the accuracy gate of every utterance of the batched output (identical
    V/UV, relative F0 error <= 1e-6) with the reference benchmark's
    accuracy figures;
  * config2: one 60 s signal at 48 kHz through the plain call, host input
    and HBM-resident input, with its kernel breakdown;
  * config5: --long recordings of 30 min at 44.1 kHz through harvest_stream,
    recording 0 checked against WORLD's stored F0.

Usage: python bench.py [--gpus 1] [--steps 10] [--warmup 5]
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import statistics
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "harvest-cuda_amd"))
sys.path.insert(0, str(ROOT))

METRIC = "audio-sec/sec on 1 MI355X, batched vs plain, same V/UV, ≤1e-6 rel F0 err"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
KERNEL_KEYS = ["kernel_decimate", "kernel_filterbank", "kernel_interp", "kernel_refine",
               "kernel_prune_compact", "kernel_refine_dft"]
WORK_KEYS = {"window_samples": "work_window_samples", "bin_samples": "work_bin_samples",
             "table_pairs": "work_table_pairs"}
KERNEL_NAMES = {"kernel_refine": "k_refine_waves", "kernel_filterbank": "k_filter_fft+k_filter_events",
                "kernel_interp": "k_tile_scan+k_interp", "kernel_decimate": "k_iir_*+k_carry_*",
                "kernel_prune_compact": "k_prune_best+scan+k_compact"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(lengths, fs, floor=71.0, ceil=800.0):
    """Implementation-independent HBM bytes per utterance, summed:
    A = 8 L + 8 N + 16 (n/2 + 1)(2B + 1) + 16 B F + 24 F C (read x, write y,
    the signal spectrum and one round trip of each band's product spectrum at
    WORLD's global FFT size n, the B x F raw candidate matrix, and the F x C
    candidate slots read and refined; C = 7 round(B / 10))."""
    from harvest_cuda import _core
    total = 0.0
    for L in lengths:
        p = _core._params(L, fs, floor, ceil, 5.0)
        n, B, F, C = p["fft_size"], p["bands"], p["frames"], 7 * p["capacity"]
        total += 8 * L + 8 * p["y_length"] + 16 * (n // 2 + 1) * (2 * B + 1) + 16 * B * F + 24 * F * C
    return total


def measured_traffic(kernel, workload):
    """HBM bytes per dispatch of `kernel` on `workload` from the committed
    FETCH_SIZE/WRITE_SIZE counter passes (profiles/*traffic*.md, each line
    `traffic_bytes_per_launch <kernel> <workload>: <bytes>`); None when not
    profiled."""
    import re
    pat = re.compile(r"^traffic_bytes_per_launch %s %s: (\d+)$" % (re.escape(kernel),
                                                                     re.escape(workload)), re.M)
    for path in sorted((ROOT / "profiles").glob("*traffic*.md")):
        m = pat.search(path.read_text())
        if m:
            return float(m.group(1))
    return None


def stage_bytes(lengths, fs, floor=71.0, ceil=800.0):
    """The per-stage terms of algorithmic_bytes(), summed over utterances:
    the filterbank's 16 (n/2 + 1)(2B + 1) (signal spectrum and one round trip
    of each band's product spectrum at WORLD's global n) and the
    refinement's 24 F C (each candidate slot read, its refined f0 and score
    written)."""
    from harvest_cuda import _core
    fb = rf = 0.0
    for L in lengths:
        p = _core._params(L, fs, floor, ceil, 5.0)
        n, B, F, C = p["fft_size"], p["bands"], p["frames"], 7 * p["capacity"]
        fb += 16 * (n // 2 + 1) * (2 * B + 1)
        rf += 24 * F * C
    return {"kernel_filterbank": fb, "kernel_refine": rf}


def filterbank_flops(lengths, fs, fft_min_half, floor=71.0, ceil=800.0):
    """FP64 FLOPs of the filterbank as this build computes it, summed over
    utterances: bands whose half length h is at least fft_min_half run by
    overlap-save FFT (per 2046-sample tile one forward real 4096-point
    transform, 2.5 n log2 n with n = 4096, then per band the spectrum
    product, 6 FLOP per bin for 2049 bins, and the inverse transform), the
    others by the direct FIR (2 FLOP per tap per output sample)."""
    import math
    from harvest_cuda import _core
    p = _core._params(lengths[0], fs, floor, ceil, 5.0)
    fs_w = p["fs_work"]
    halves = []
    for b in range(p["bands"]):
        x = fs_w / (floor * 0.9 * 2.0 ** ((b + 1) / 40.0)) * 2.0
        halves.append(int(x + 0.5))
    nfft = 0
    if halves[0] <= 1000:
        while nfft < len(halves) and fft_min_half <= halves[nfft] <= halves[0]:
            nfft += 1
    taps = sum(2 * h + 1 for h in halves[nfft:])
    fft4096 = 2.5 * 4096 * 12
    total = 0.0
    for L in lengths:
        n = _core._params(L, fs, floor, ceil, 5.0)["y_length"]
        tiles = math.ceil(n / 2046)
        total += 2.0 * taps * n
        if nfft:
            total += tiles * fft4096 + tiles * nfft * (6 * 2049 + fft4096)
    return total


def roofline(kernels_ms, work, launches, workload, lengths, fs):
    """Roofline objects of the refinement's DFT kernel and of the filterbank,
    and of the one with the larger time per step.  Both are bound by FP64
    vector FMAs (no MFMA: neither is a dense contraction), against the
    78.6 TFLOP/s FP64 vector peak (MI355X's FP64 matrix peak is the same).
    At working rates that are multiples of 1000 Hz the refinement runs the
    bin-table path, whose DFT kernel k_bt_dft computes, per (frame, half
    window) pair, every bin of that half window's table by a Goertzel
    recurrence over the main- and derivative-windowed samples: 6 FLOP per
    bin-sample (an FMA and a subtraction per signal, an FMA counted as 2)
    plus 2 per window sample (the two window products), counted on the
    device; it is timed alone (its time includes FixF0 of the slots that
    read each pair, whose FLOPs are not counted).  Other rates run k_refine_waves: 8 FLOP per
    bin-sample plus 17 per window sample, timed with its small set-up
    kernels.  The filterbank's FLOPs are filterbank_flops().  `launches` is
    the analyses per step (one per bucket): a launch here is one analysis'
    refinement DFT kernel, or its filterbank stage (tile spectra, FFT bands,
    direct bands).  Each object also carries `hbm_view`: the same launch
    time against HBM bandwidth, by the stage's term of the byte count
    algorithmic_bytes() uses (stage_bytes())."""
    from harvest_cuda import tuning
    fft_min = tuning.current()["backend"]["fft_min_half"]
    out = {}
    if work["table_pairs"] > 0:
        per = kernels_ms["kernel_refine_dft"] / launches
        flops = (6.0 * work["bin_samples"] + 2.0 * work["window_samples"]) / launches
        kernel, model = "k_bt_dft", "6 FLOP per bin-sample + 2 FLOP per window sample"
    else:
        per = kernels_ms["kernel_refine"] / launches
        flops = (8.0 * work["bin_samples"] + 17.0 * work["window_samples"]) / launches
        kernel, model = "k_refine_waves", "8 FLOP per bin-sample + 17 FLOP per window sample"
    ach = flops / (per / 1000.0) / 1e12
    out["kernel_refine"] = {
        "bound": "fp64_valu", "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
        "frac": ach / FP64_PEAK_TFLOPS, "traffic": measured_traffic(kernel, workload),
        "kernel": kernel, "avg_launch_ms": per, "launches_per_step": launches,
        "flop_per_launch": flops, "bin_samples_per_launch": work["bin_samples"] / launches,
        "window_samples_per_launch": work["window_samples"] / launches,
        "refinement_stage_ms_per_launch": kernels_ms["kernel_refine"] / launches,
        "model": model}
    per = kernels_ms["kernel_filterbank"] / launches
    flops = filterbank_flops(lengths, fs, fft_min) / launches
    ach = flops / (per / 1000.0) / 1e12
    out["kernel_filterbank"] = {
        "bound": "fp64_valu", "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
        "frac": ach / FP64_PEAK_TFLOPS,
        "traffic": measured_traffic("k_spectrum+k_filter_fft+k_filter_events", workload),
        "kernel": "k_spectrum+k_filter_fft+k_filter_events", "avg_launch_ms": per,
        "launches_per_step": launches, "flop_per_launch": flops,
        "model": f"FFT bands (h >= {fft_min}): 2.5 n log2 n per 4096-point transform + 6 FLOP "
                 "per bin; direct bands: 2 FLOP per tap per output sample"}
    # The same launches against HBM, by the stage's term of the
    # implementation-independent byte count (algorithmic_bytes()).
    terms = stage_bytes(lengths, fs)
    for k, r in out.items():
        per_bytes = terms[k] / launches
        gbps = per_bytes / (r["avg_launch_ms"] / 1000.0) / 1e9
        r["hbm_view"] = {"algorithmic_bytes_per_launch": per_bytes, "achieved_gbps": gbps,
                         "peak_gbps": HBM_PEAK_GBS, "frac": gbps / HBM_PEAK_GBS,
                         "model": "16(n/2+1)(2B+1) per utterance" if k == "kernel_filterbank"
                                  else "24FC per utterance"}
    name = max(("kernel_refine", "kernel_filterbank"),
               key=lambda k: out[k]["avg_launch_ms"])
    return out[name], out


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def host_threads():
    """The host's cores, at most 16 (the share of one GPU's host on the box)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def world_over(xs, fs, threads):
    """WORLD Harvest of every utterance on `threads` host threads, longest
    first (each call releases the GIL); returns (f0 list, seconds, kind)."""
    from oracle import oracle
    kind, fn = ("reference", oracle.ref_harvest) if oracle.ref_available() else ("port", oracle.harvest)
    order = sorted(range(len(xs)), key=lambda i: -len(xs[i]))
    out = [None] * len(xs)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as pool:
        for i, (f0, _) in zip(order, pool.map(lambda i: fn(xs[i], fs), order)):
            out[i] = f0
    return out, time.perf_counter() - t0, kind


def accuracy(expected, actual):
    """The acceptance bar over many tracks, and the reference benchmark's
    figures (harvest_cuda.metrics.compare) over their concatenation."""
    import numpy as np
    from harvest_cuda.metrics import compare
    vuv_bad, worst, bad_utts = 0, 0.0, 0
    for e, a in zip(expected, actual):
        v = e > 0
        mism = int(np.count_nonzero(v != (a > 0)))
        both = v & (a > 0)
        rel = float(np.max(np.abs(a[both] - e[both]) / e[both])) if both.any() else 0.0
        vuv_bad += mism
        worst = max(worst, rel)
        bad_utts += mism > 0 or rel > 1e-6
    acc = compare(np.concatenate(expected), np.concatenate(actual))
    return {"utterances": len(expected), "utterances_failing": bad_utts,
            "vuv_identical": vuv_bad == 0, "vuv_mismatches": vuv_bad, "max_rel_err": worst,
            "within_1e-6": bad_utts == 0, "compare": acc}


def batched_leg(args, result):
    import numpy as np
    import harvest_cuda
    from harvest_cuda.signals import config3
    t0 = time.perf_counter()
    xs, fs = config3()
    lengths = [len(x) for x in xs]
    audio = sum(lengths) / fs
    log(f"config 3: {len(xs)} utterances, {audio:.1f} audio-s, generated in "
        f"{time.perf_counter() - t0:.1f} s")

    t0 = time.perf_counter()
    harvest_cuda.harvest_batch(xs, fs)
    first_ms = (time.perf_counter() - t0) * 1000.0
    for _ in range(args.warmup):  # with timings, so that their events exist before timing
        harvest_cuda._core.harvest_batch(xs, fs, timings=True)
    # The interpreter's own objects (torch's module graph among them) are
    # moved out of the collector's generations, so that a full collection in
    # the timed loop does not traverse them (about 35 ms otherwise).
    gc.collect()
    gc.freeze()
    kernels = {k: 0.0 for k in KERNEL_KEYS}
    work = {k: 0.0 for k in WORK_KEYS}
    device_ms, launches = 0.0, 0
    steps_ms = []
    t0 = time.perf_counter()
    last = t0
    for _ in range(args.steps):
        out, tm = harvest_cuda._core.harvest_batch(xs, fs, timings=True)
        now = time.perf_counter()
        steps_ms.append((now - last) * 1000.0)
        log(f"  step {steps_ms[-1]:.1f} ms: call {tm.get('call', 0):.1f}, host submit "
            f"{tm.get('host_submit', 0):.1f}, wait {tm.get('host_wait', 0):.1f} (max "
            f"{tm.get('host_wait_max', 0):.1f}), contour {tm.get('host_contour', 0):.1f}")
        last = now
        for k in KERNEL_KEYS:
            kernels[k] += tm[k]
        for k, key in WORK_KEYS.items():
            work[k] += tm[key]
        device_ms += tm["preprocess"] + tm["filter_events"] + tm["refine"]
        launches += int(tm.get("groups", 1))
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / args.steps * 1000.0
    log("batched step ms: " + " ".join(f"{v:.1f}" for v in steps_ms))
    batched = out

    # Plain: one harvest() per utterance, same scope.
    plain_ms = []
    plain = None
    for rep in range(args.plain_repeats + 1 if args.plain_repeats > 0 else 0):
        # the first loop warms every shape's buffers
        t1 = time.perf_counter()
        plain = [harvest_cuda.harvest(x, fs) for x in xs]
        if rep:
            plain_ms.append((time.perf_counter() - t1) * 1000.0)
    identical = plain is not None and all(
        np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) for a, b in zip(batched, plain))
    plain_med = statistics.median(plain_ms) if plain_ms else None
    log(f"batched {ms_per_step:.1f} ms/step ({audio / ms_per_step * 1000:.0f} audio-s/s), plain "
        f"{plain_med and round(plain_med, 1)} ms, bitwise equal {identical}")

    steps = args.steps
    workload = "config3"
    result.update({
        "value": audio / (ms_per_step / 1000.0), "ms_per_step": ms_per_step,
        "config": {"workload": "512 float64 utterances at 24 kHz, lengths log-uniform 0.5-30 s "
                               "(seed 1003), one batched call, NumPy in and out",
                   "fs": fs, "utterances": len(xs), "audio_seconds": audio,
                   "f0_floor": 71.0, "f0_ceil": 800.0, "frame_period": 5.0},
        "step_ms": {"median": statistics.median(steps_ms), "min": min(steps_ms),
                    "max": max(steps_ms)},
        "first_call_ms": first_ms,
        "plain": {"ms_per_loop_median": plain_med, "loops": len(plain_ms),
                  "audio_sec_per_sec": audio / (plain_med / 1000.0) if plain_med else None,
                  "calls_per_loop": len(xs),
                  "note": "one harvest(x) call per utterance, NumPy in and out"},
        "batched_over_plain": (plain_med / ms_per_step) if plain_med else None,
        "batched_equals_plain_bitwise": bool(identical) if plain_ms else None,
        "kernels_ms_per_step": {k: v / steps for k, v in kernels.items()},
        "device_ms_per_step": device_ms / steps,
    })
    result["roofline"], result["rooflines"] = roofline(
        {k: v / steps for k, v in kernels.items()}, {k: v / steps for k, v in work.items()},
        launches // steps, workload, lengths, fs)
    a_bytes = algorithmic_bytes(lengths, fs)
    dev_s = device_ms / steps / 1000.0
    result["algorithmic_hbm"] = {
        "bytes_per_step": a_bytes, "gbps_device_span": a_bytes / dev_s / 1e9,
        "gbps_wall": a_bytes / (ms_per_step / 1000.0) / 1e9,
        "frac_of_peak_device_span": a_bytes / dev_s / 1e9 / HBM_PEAK_GBS,
        "model": "8L + 8N + 16(n/2+1)(2B+1) + 16BF + 24FC per utterance (global-n spectra; this "
                 "design never materialises them, so the figure is a normalised throughput)"}
    if args.cpu_baseline:
        threads = host_threads()
        ref, dt, kind = world_over(xs, fs, threads)
        result["cpu_baseline"] = {
            "value": audio / dt, "unit": "audio-sec/sec", "cores": threads, "kind": kind,
            "cpu": cpu_model(), "seconds": dt,
            "sample": f"all {len(xs)} utterances ({audio:.0f} audio-s), one WORLD Harvest call "
                      f"each on a pool of {threads} threads, longest first"}
        result["parity"] = dict(accuracy(ref, [f for f, _ in batched]),
                                against="WORLD Harvest" if kind == "reference" else "oracle restatement")
        log(f"cpu baseline ({kind}, {threads} threads): {audio / dt:.1f} audio-s/s; "
            f"parity {result['parity']}")
    return xs, fs


def config2_leg(args):
    import numpy as np
    import torch
    from harvest_cuda import _core
    from harvest_cuda.signals import config2
    x, fs = config2()
    seconds = len(x) / fs
    dx = torch.from_numpy(x).to("cuda:0")
    torch.cuda.synchronize()
    ptr = dx.data_ptr()
    for _ in range(5):
        _core.harvest_device(ptr, [len(x)], fs, timings=True)
    steps = max(args.steps, 10)
    kernels = {k: 0.0 for k in KERNEL_KEYS}
    work = {k: 0.0 for k in WORK_KEYS}
    device_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        out, tm = _core.harvest_device(ptr, [len(x)], fs, timings=True)
        for k in KERNEL_KEYS:
            kernels[k] += tm[k]
        for k, key in WORK_KEYS.items():
            work[k] += tm[key]
        device_ms += tm["preprocess"] + tm["filter_events"] + tm["refine"]
    hbm_ms = (time.perf_counter() - t0) / steps * 1000.0
    for _ in range(2):
        _core.harvest(x, fs)
    t0 = time.perf_counter()
    for _ in range(steps):
        f0, _t = _core.harvest(x, fs)
    host_ms = (time.perf_counter() - t0) / steps * 1000.0
    same = bool(np.array_equal(out[0][0], f0))
    return {"workload": "one 60 s float64 signal at 48 kHz (seed 1002), plain drop-in call",
            "steps": steps, "host_input": {"ms_per_call": host_ms,
                                           "audio_sec_per_sec": seconds / (host_ms / 1000.0)},
            "hbm_input": {"ms_per_call": hbm_ms, "audio_sec_per_sec": seconds / (hbm_ms / 1000.0)},
            "device_ms": device_ms / steps,
            "kernels_ms_per_call": {k: v / steps for k, v in kernels.items()},
            "rooflines": roofline({k: v / steps for k, v in kernels.items()},
                                  {k: v / steps for k, v in work.items()}, 1, "config2",
                                  [len(x)], fs)[1],
            "host_equals_hbm_input": same}


def config5_leg(args):
    import numpy as np
    import harvest_cuda
    from harvest_cuda.signals import config5_recording
    n = args.long
    t0 = time.perf_counter()
    with ThreadPoolExecutor(n) as pool:
        recs = list(pool.map(lambda i: config5_recording(i)[0], range(n)))
    fs = 44100
    log(f"config 5: {n} recordings generated in {time.perf_counter() - t0:.1f} s")
    harvest_cuda.harvest(recs[-1], fs)  # the first call per shape allocates ~15 GB
    times, outs = [], []
    last = time.perf_counter()
    for f0, _ in harvest_cuda.harvest_stream(iter(recs), fs):
        now = time.perf_counter()
        times.append((now - last) * 1000.0)
        last = now
        outs.append(f0)
    seconds = len(recs[0]) / fs
    a_bytes = algorithmic_bytes([len(recs[0])], fs)
    res = {"workload": f"{n} recordings of 30 min at 44.1 kHz (seed 1005) through harvest_stream",
           "ms_per_recording": times, "audio_sec_per_sec": n * seconds / (sum(times) / 1000.0),
           "algorithmic_gbps": a_bytes * n / (sum(times) / 1000.0) / 1e9}
    golden = ROOT / "tests" / "golden" / "world_long.npz"
    if golden.exists():
        import hashlib
        g = np.load(golden)
        if hashlib.sha256(recs[0].tobytes()).hexdigest() == str(g["sha"]):
            ref = g["f0"].astype(np.float64)
            res["parity_recording0"] = dict(accuracy([ref], [outs[0]]), against="WORLD Harvest "
                                            "(stored F0, float32)")
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--plain-repeats", type=int, default=3)
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--batched", type=int, default=1,
                    help="run the batched leg (0: skip it, to profile the other legs alone)")
    ap.add_argument("--config2", type=int, default=1, help="also run the 60 s plain-call leg")
    ap.add_argument("--long", type=int, default=2, help="30-min recordings of the long leg (0: skip)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-MiB float64 tensor 5 times on the GPU (a kernel of known "
                         "traffic, to calibrate the profiler's FETCH_SIZE/WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 8
        src = torch.rand(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "launches": 5}))
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(utterances would shard across GPUs with no exchange); run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    if args.steps < 1:
        raise SystemExit("--steps must be at least 1")

    import harvest_cuda
    if not harvest_cuda.cuda_available(0):
        raise SystemExit("no supported gfx950 GPU visible")
    result = {"metric": METRIC, "value": None, "unit": "audio-sec/sec", "n_gpus": 1,
              "steps": args.steps, "warmup": args.warmup, "ms_per_step": None,
              "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
              "data": "synthetic"}
    if args.batched:
        batched_leg(args, result)
    if args.config2:
        result["config2"] = config2_leg(args)
        log(f"config 2: {result['config2']['host_input']}")
    if args.long > 0:
        result["config5"] = config5_leg(args)
        log(f"config 5: {result['config5']}")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
