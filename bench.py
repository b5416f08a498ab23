#!/usr/bin/env python
"""Throughput benchmark of the MI355X Harvest path.

Primary line (BASELINE metric): audio-seconds per second of the plain drop-in
call on the config-2 workload (one 60 s float64 signal at 48 kHz), input
already resident in HBM, output f0 on the host (device analysis + host
contour).  One step = one call.  Also reported:
  * roofline of the dominant kernel from hipEvents recorded around every
    kernel launch on the backend's stream during the timed steps;
  * parity of the step's output against WORLD (same V/UV, max relative F0
    error on voiced frames);
  * cpu_baseline: WORLD Harvest compiled from the reference (oracle/_ref) or,
    where that library is absent, the oracle's C restatement, single thread,
    on the same 60 s signal;
  * batched: config 3 (512 utterances at 24 kHz) as one batched call versus
    one plain call per utterance on the same GPU.

Usage: python bench.py [--gpus 1] [--steps 50] [--warmup 40] [--batched 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "harvest-cuda_amd"))
sys.path.insert(0, str(ROOT))

METRIC = "audio-sec/sec on 1 MI355X, batched vs plain, same V/UV, ≤1e-6 rel F0 err"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def fir_flops(lengths, fs, floor=71.0, ceil=800.0):
    """Useful FP64 FLOPs of the filterbank kernel: one FMA (2 FLOPs) per tap
    per output sample per band, 2 sum(2h_b + 1) N_u over utterances."""
    from harvest_cuda import _core
    import math
    p = _core._params(lengths[0], fs, floor, ceil, 5.0)
    fs_w = p["fs_work"]
    af = floor * 0.9
    taps = 0
    for b in range(p["bands"]):
        bound = af * 2.0 ** ((b + 1) / 40.0)
        x = fs_w / bound * 2.0
        h = int(x + 0.5) if x > 0 else int(x - 0.5)
        taps += 2 * h + 1
    n_total = sum(_core._params(n, fs, floor, ceil, 5.0)["y_length"] for n in lengths)
    return 2.0 * taps * n_total


def algorithmic_bytes(lengths, fs, floor=71.0, ceil=800.0):
    """Implementation-independent HBM bytes of the path per call, summed over
    utterances: A = 8 L + 8 N + 16 (n/2 + 1)(2B + 1) + 16 B F + 24 F C (read x,
    write y, the signal spectrum and one round trip of each band's product
    spectrum at WORLD's global FFT size n, the B x F raw candidate matrix, and
    the F x C candidate slots read and refined; C = 7 round(B / 10))."""
    from harvest_cuda import _core
    total = 0.0
    for L in lengths:
        p = _core._params(L, fs, floor, ceil, 5.0)
        n, B, F, C = p["fft_size"], p["bands"], p["frames"], 7 * p["capacity"]
        total += 8 * L + 8 * p["y_length"] + 16 * (n // 2 + 1) * (2 * B + 1) + 16 * B * F + 24 * F * C
    return total


def measured_traffic(kernel):
    """HBM bytes per dispatch of `kernel` on this workload, from the committed
    FETCH_SIZE/WRITE_SIZE counter passes calibrated by a copy of known size
    (profiles/config2_traffic.md); None when not profiled."""
    import re
    path = ROOT / "profiles" / "config2_traffic.md"
    if not path.exists():
        return None
    m = re.search(r"^traffic_bytes_per_launch %s: (\d+)$" % re.escape(kernel), path.read_text(), re.M)
    return float(m.group(1)) if m else None


def roofline(kernels_ms, work, lengths, fs, steps):
    """Roofline object of the kernel with the largest average duration."""
    avg = {k: v / steps for k, v in kernels_ms.items()}
    name = max(avg, key=avg.get)
    sec = avg[name] / 1000.0
    if name == "kernel_filterbank":
        flops = fir_flops(lengths, fs)
        achieved = flops / sec / 1e12
        return name, {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                      "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                      "traffic": measured_traffic("k_filter_events"),
                      "kernel": "k_filter_events", "avg_ms": avg[name],
                      "model": "2 FLOP per tap per output sample per band (FP64 FMA)"}
    if name == "kernel_refine":
        flops = 8.0 * work["bin_samples"] / steps + 17.0 * work["window_samples"] / steps
        achieved = flops / sec / 1e12
        return name, {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                      "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                      "traffic": measured_traffic("k_refine_waves"),
                      "kernel": "k_refine_waves", "avg_ms": avg[name],
                      "model": "8 FLOP per bin-sample (4 FMA) + 17 FLOP per window sample, "
                               "cosines not counted"}
    # HBM-bound stages: algorithmic bytes of the decimation scan (read x
    # twice, write and read the pass buffer twice, write y).
    total_x = sum(lengths)
    bytes_ = 8.0 * (2 * total_x + 4 * total_x + total_x / 3)
    achieved = bytes_ / sec / 1e9
    return name, {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": name,
                  "avg_ms": avg[name]}


def world_baseline(x, fs):
    from oracle import oracle
    if oracle.ref_available():
        kind, fn = "reference", oracle.ref_harvest
    else:
        kind, fn = "port", oracle.harvest
    t0 = time.perf_counter()
    f0, _ = fn(x, fs)
    dt = time.perf_counter() - t0
    return f0, dt, kind


def parity(actual, expected):
    import numpy as np
    v = expected > 0
    both = v & (actual > 0)
    rel = float(np.max(np.abs(actual[both] - expected[both]) / expected[both])) if both.any() else 0.0
    return {"vuv_identical": bool(np.array_equal(actual > 0, v)),
            "vuv_mismatches": int(np.count_nonzero((actual > 0) != v)),
            "max_rel_err": rel, "voiced_frames": int(v.sum())}


def batched_leg(device, warmup):
    import numpy as np
    import torch
    from harvest_cuda import _core
    from harvest_cuda.signals import config3
    t0 = time.perf_counter()
    xs, fs = config3()
    log(f"config 3 generated in {time.perf_counter() - t0:.1f} s")
    lengths = [len(x) for x in xs]
    audio = sum(lengths) / fs
    flat = torch.from_numpy(np.concatenate(xs)).to(device)
    torch.cuda.synchronize()
    base = flat.data_ptr()
    offs = np.concatenate([[0], np.cumsum(lengths)])[:-1]

    def batched():
        return _core.harvest_device(base, lengths, fs)

    def plain():
        return [_core.harvest_device(base + 8 * int(o), [n], fs)[0] for o, n in zip(offs, lengths)]

    for _ in range(warmup):
        batched()
    plain()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rb = batched()
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    t0 = time.perf_counter()
    rp = plain()
    torch.cuda.synchronize()
    tp = time.perf_counter() - t0
    identical = all(np.array_equal(a[0], b[0]) for a, b in zip(rb, rp))
    return {"workload": "config 3: 512 utterances at 24 kHz, lengths log-uniform 0.5-30 s",
            "audio_seconds": audio, "batched_audio_sec_per_sec": audio / tb,
            "plain_audio_sec_per_sec": audio / tp, "speedup": tp / tb,
            "batched_ms": tb * 1000, "plain_ms": tp * 1000,
            "batched_equals_plain": bool(identical)}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    # A process's first calls include one-time costs: one stall of 6-9 ms on
    # one of roughly the first 30 calls (not tied to the number of launches,
    # copies or events on the analysis stream: warming those at workspace
    # creation does not move it).  The warm-up absorbs it.
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--batched", type=int, default=1, help="also run the config-3 batched leg")
    ap.add_argument("--cpu-baseline", type=int, default=1)
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

    import numpy as np
    import torch
    import harvest_cuda
    from harvest_cuda import _core
    from harvest_cuda.signals import config2

    if not harvest_cuda.cuda_available(0):
        raise SystemExit("no supported gfx950 GPU visible")
    device = torch.device("cuda:0")
    x, fs = config2()
    seconds = len(x) / fs
    dx = torch.from_numpy(x).to(device)
    torch.cuda.synchronize()
    ptr = dx.data_ptr()

    def step(timed):
        return _core.harvest_device(ptr, [len(x)], fs, timings=timed)

    for _ in range(args.warmup):  # with timings, so that their events exist before timing
        step(True)
    kernel_keys = ["kernel_decimate", "kernel_filterbank", "kernel_interp", "kernel_refine",
                   "kernel_prune_compact"]
    kernels = {k: 0.0 for k in kernel_keys}
    work = {"window_samples": 0.0, "bin_samples": 0.0}
    device_ms = 0.0  # hipEvent span of the device launch sequence
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    marks = []
    for _ in range(args.steps):
        out, tm = step(True)
        marks.append(time.perf_counter())
        for k in kernel_keys:
            kernels[k] += tm[k]
        work["window_samples"] += tm["work_window_samples"]
        work["bin_samples"] += tm["work_bin_samples"]
        device_ms += tm["preprocess"] + tm["filter_events"] + tm["refine"]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / args.steps * 1000.0
    per = [(b - a) * 1000.0 for a, b in zip([t0] + marks[:-1], marks)]
    log("step ms: " + " ".join(f"{v:.2f}" for v in per))
    f0 = out[0][0]
    name, roof = roofline(kernels, work, [len(x)], fs, args.steps)
    log(f"plain config 2: {ms_per_step:.3f} ms/step, dominant {name} {roof['avg_ms']:.3f} ms")

    result = {
        "metric": METRIC, "value": seconds / (ms_per_step / 1000.0), "unit": "audio-sec/sec",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "config 2: one 60 s float64 signal at 48 kHz, plain drop-in call, "
                               "input resident in HBM, f0 returned to the host",
                   "fs": fs, "seconds": seconds, "f0_floor": 71.0, "f0_ceil": 800.0,
                   "frame_period": 5.0},
        "roofline": roof,
        "kernels_ms_per_step": {k: v / args.steps for k, v in kernels.items()},
    }
    # The same call with the waveform on the host (NumPy in, PCIe upload
    # included): the drop-in API's rate, reported beside the primary value.
    for _ in range(2):
        _core.harvest(x, fs)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _core.harvest(x, fs)
    host_ms = (time.perf_counter() - t0) / args.steps * 1000.0
    result["host_input"] = {"ms_per_step": host_ms, "audio_sec_per_sec": seconds / (host_ms / 1000.0),
                            "note": "harvest(x) with x a host NumPy array: includes the H2D copy"}
    a_bytes = algorithmic_bytes([len(x)], fs)
    dev_s = device_ms / args.steps / 1000.0
    result["algorithmic_hbm"] = {
        "bytes_per_step": a_bytes, "device_ms": dev_s * 1000.0,
        "gbps_device_span": a_bytes / dev_s / 1e9, "gbps_wall": a_bytes / (ms_per_step / 1000.0) / 1e9,
        "frac_of_peak_device_span": a_bytes / dev_s / 1e9 / HBM_PEAK_GBS,
        "model": "8L + 8N + 16(n/2+1)(2B+1) + 16BF + 24FC (global-n spectra; this design "
                 "never materialises them, so the figure is a normalised throughput)"}
    if args.cpu_baseline:
        ref_f0, dt, kind = world_baseline(x, fs)
        result["cpu_baseline"] = {"value": seconds / dt, "unit": "audio-sec/sec", "cores": 1,
                                  "kind": kind,
                                  "sample": "the full config-2 signal (60 s at 48 kHz), one call, "
                                            "single thread (WORLD Harvest is single-threaded)"}
        result["parity"] = dict(parity(f0, ref_f0), against="WORLD Harvest" if kind == "reference"
                                else "oracle restatement")
        log(f"cpu baseline ({kind}): {seconds / dt:.2f} audio-s/s; parity {result['parity']}")
    if args.batched:
        result["batched"] = batched_leg(device, 1)
        log(f"batched: {result['batched']}")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
