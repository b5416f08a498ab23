"""Autotuner of the batched form (run on an MI355X):

    python -m harvest_cuda.autotune [--out harvest_cuda/tuning.json]

It measures, on a calibration workload drawn like the batched benchmark
(512 utterances at 24 kHz, lengths log-uniform over 0.5-30 s) but from
another seed (2003), so that the benchmark's own utterances are not tuned
on:
  1. the FFT/direct filterbank split (fft_min_half): the filterbank kernels'
     time on one 60 s utterance at 24 kHz for each candidate split, keeping
     only splits whose F0 equals the default's to 1e-9 with identical
     voicing; then, on the same utterance, the persistent grid of the
     bin-table refinement's DFT kernel (refine_groups), by that kernel's
     time, keeping only grids whose F0 equals the default's bit for bit;
  2. the pipeline: how many length buckets (max_groups) and how many
     workspaces keep buckets in flight (depth 2-4), jointly, from the wall
     time of one batched call, NumPy in and out, median of 3; more buckets
     in flight overlap one bucket's host staging, copies and contours with
     the others' kernels, at the price of more, smaller launches;
  3. the bucket boundaries (first_weight, last_weight) at that pipeline;
  4. whether the buckets' device work runs one after another (serialize);
  5. the crossing slabs' memory budget per band group (slab_budget_mib),
     which sets how many bands one filterbank launch covers;
  6. HIP graph capture on or off at the chosen pipeline.
The winners and every measurement go to the table, which the package
applies at import (harvest_cuda/tuning.py).
"""
from __future__ import annotations

import argparse
import itertools
import json
import statistics
import sys
import time
from pathlib import Path

import numpy as np

from . import _core, tuning
from .signals import _voiced_unvoiced, config3

FFT_SPLITS = (32, 48, 64, 80, 96, 128, 160, 1001)
REFINE_GROUPS = (512, 768, 1024, 1536, 2048, 3072)
SLAB_BUDGETS = (1024, 2048, 4096, 8192)
DEPTHS = (2, 3, 4)
GROUPS = (4, 8, 12, 16, 24, 32)
WEIGHTS = (0.4, 0.7, 1.0)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def calibration_utterance():
    """One 60 s utterance at 24 kHz in device memory: (tensor, length, fs)."""
    import torch
    rng = np.random.default_rng(np.random.SeedSequence(2003).spawn(1)[0])
    fs = 24000
    x = np.ascontiguousarray(_voiced_unvoiced(rng, 60 * fs, fs, 25.0))
    dx = torch.from_numpy(x).to("cuda:0")
    torch.cuda.synchronize()
    return dx, len(x), fs


def tune_fft_split(repeats: int = 10) -> tuple[int, list]:
    dx, n, fs = calibration_utterance()
    rows, ref = [], None
    for split in FFT_SPLITS:
        tuning.set_backend(fft_min_half=split)
        for _ in range(3):
            _core.harvest_device(dx.data_ptr(), [n], fs, timings=True)
        ms = []
        for _ in range(repeats):
            out, tm = _core.harvest_device(dx.data_ptr(), [n], fs, timings=True)
            ms.append(tm["kernel_filterbank"])
        f0 = out[0][0]
        if ref is None and split == 64:
            ref = f0
        rows.append({"fft_min_half": split, "filterbank_ms": statistics.median(ms), "f0": f0})
        log(f"fft_min_half {split}: filterbank {rows[-1]['filterbank_ms']:.4f} ms")
    ok = []
    for r in rows:
        f0 = r.pop("f0")
        v = ref > 0
        same_vuv = bool(np.array_equal(v, f0 > 0))
        rel = float(np.max(np.abs(f0[v] - ref[v]) / ref[v])) if v.any() else 0.0
        r.update(same_vuv=same_vuv, max_rel_vs_default=rel)
        if same_vuv and rel <= 1e-9:
            ok.append(r)
    best = min(ok, key=lambda r: r["filterbank_ms"])["fft_min_half"]
    return best, rows


def tune_refine_groups(repeats: int = 10) -> tuple[int, list]:
    dx, n, fs = calibration_utterance()
    rows, ref = [], None
    for groups in (0,) + REFINE_GROUPS:  # 0: the backend's default, the reference output
        tuning.set_backend(refine_groups=groups)
        for _ in range(3):
            _core.harvest_device(dx.data_ptr(), [n], fs, timings=True)
        ms = []
        for _ in range(repeats):
            out, tm = _core.harvest_device(dx.data_ptr(), [n], fs, timings=True)
            ms.append(tm["kernel_refine_dft"])
        f0 = out[0][0]
        if ref is None:
            ref = f0
        rows.append({"refine_groups": groups, "dft_ms": statistics.median(ms),
                     "same_f0": bool(np.array_equal(f0, ref))})
        log(f"refine_groups {groups}: k_bt_dft {rows[-1]['dft_ms']:.4f} ms")
    ok = [r for r in rows if r["same_f0"] and r["refine_groups"] > 0]
    best = min(ok, key=lambda r: r["dft_ms"])["refine_groups"] if ok else 0
    return best, rows


def batch_wall(xs, fs, repeats: int) -> float:
    _core.harvest_batch(xs, fs)
    ms = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        _core.harvest_batch(xs, fs)
        ms.append((time.perf_counter() - t0) * 1000.0)
    return statistics.median(ms)


def tune_pipeline(xs, fs, repeats: int = 3) -> tuple[dict, list]:
    base = {"min_group_frames": 20000, "split_frames": 200000}
    rows = []

    def measure(**kw):
        tuning.set_batch(**base, **kw)
        ms = batch_wall(xs, fs, repeats)
        rows.append({**kw, "ms": ms})
        log(" ".join(f"{k} {v}" for k, v in kw.items()) + f": {ms:.1f} ms")
        return ms

    chosen = {"first_weight": 0.4, "last_weight": 0.4, "serialize": 0}
    best = None
    for depth, g in itertools.product(DEPTHS, GROUPS):
        ms = measure(max_groups=g, depth=depth, **chosen)
        if best is None or ms < best[0]:
            best = (ms, depth, g)
    chosen.update(depth=best[1], max_groups=best[2])
    for first, last in itertools.product(WEIGHTS, WEIGHTS):
        ms = measure(**{**chosen, "first_weight": first, "last_weight": last})
        if ms < best[0]:
            best = (ms,) + best[1:]
            chosen.update(first_weight=first, last_weight=last)
    ms = measure(**{**chosen, "serialize": 1})
    if ms < best[0]:
        chosen["serialize"] = 1
    return {**base, **chosen}, rows


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", default=str(tuning.TABLE))
    ap.add_argument("--count", type=int, default=512)
    args = ap.parse_args()
    if not _core.cuda_available(0):
        raise SystemExit("no supported gfx950 GPU visible")
    t0 = time.perf_counter()
    tuning.set_backend(fft_min_half=64, graphs=0, slab_budget_mib=4096, refine_groups=0)
    split, fft_rows = tune_fft_split()
    tuning.set_backend(fft_min_half=split)
    groups, group_rows = tune_refine_groups()
    tuning.set_backend(refine_groups=groups)
    xs, fs = config3(seed=2003, count=args.count)
    audio = sum(len(x) for x in xs) / fs
    log(f"calibration batch: {len(xs)} utterances, {audio:.0f} audio-s")
    batch, pipeline_rows = tune_pipeline(xs, fs)
    tuning.set_batch(**batch)
    slab_rows = []
    for budget in SLAB_BUDGETS:
        tuning.set_backend(slab_budget_mib=budget)
        ms = batch_wall(xs, fs, 5)
        slab_rows.append({"slab_budget_mib": budget, "ms": ms})
        log(f"slab_budget_mib {budget}: {ms:.1f} ms")
    slab = min(slab_rows, key=lambda r: r["ms"])["slab_budget_mib"]
    tuning.set_backend(slab_budget_mib=slab)
    graph_rows = []
    for graphs in (0, 1):
        tuning.set_backend(graphs=graphs)
        ms = batch_wall(xs, fs, 5)
        graph_rows.append({"graphs": graphs, "ms": ms})
        log(f"graphs {graphs}: {ms:.1f} ms")
    graphs = min(graph_rows, key=lambda r: r["ms"])["graphs"]
    table = {
        "backend": {"fft_min_half": split, "graphs": graphs, "slab_budget_mib": slab,
                    "refine_table": 1, "refine_groups": groups},
        "batch": batch,
        "calibration": {"utterances": len(xs), "audio_seconds": audio, "fs": fs, "seed": 2003,
                        "seconds_to_tune": time.perf_counter() - t0},
        "measured": {"fft_split": fft_rows, "refine_groups": group_rows, "pipeline": pipeline_rows,
                     "slab_budget": slab_rows, "graphs": graph_rows},
    }
    Path(args.out).write_text(json.dumps(table, indent=1) + "\n")
    log(f"wrote {args.out}: backend {table['backend']}, batch {batch}")
    print(json.dumps({"backend": table["backend"], "batch": batch}))


if __name__ == "__main__":
    main()
