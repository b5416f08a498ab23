"""The autotuned settings of the batched form (harvest_cuda/tuning.json).

``tuning.json`` holds what ``harvest_cuda.autotune`` chose on an MI355X:
  backend: fft_min_half (FFT/direct filterbank split), graphs (HIP graph
           capture of the launch sequence), slab_budget_mib, refine_table
           (bin-table refinement where frames start at whole samples);
  batch:   how large batches are cut into length buckets (max_groups,
           min_group_frames, split_frames, first_weight, last_weight) and
           whether the buckets' device work runs one after another
           (serialize), and how many workspaces hold buckets in flight
           (depth);
and the measurements behind the choice.  ``apply()`` runs at import.
"""
from __future__ import annotations

import json
from pathlib import Path

from . import _core

TABLE = Path(__file__).resolve().parent / "tuning.json"
BACKEND_KEYS = ("fft_min_half", "graphs", "slab_budget_mib", "refine_table", "refine_groups")
BATCH_KEYS = ("max_groups", "min_group_frames", "split_frames", "first_weight", "last_weight",
              "serialize", "depth")


def load(path: Path = TABLE) -> dict:
    return json.loads(path.read_text()) if path.exists() else {}


def set_backend(**kw) -> None:
    cur = dict(_core._get_tuning())
    cur.update(kw)
    _core._set_tuning(**{k: int(cur[k]) for k in BACKEND_KEYS})


def set_batch(**kw) -> None:
    cur = dict(_core._get_batch_tuning())
    cur.update(kw)
    _core._set_batch_tuning(int(cur["max_groups"]), int(cur["min_group_frames"]),
                            int(cur["split_frames"]), float(cur["first_weight"]),
                            float(cur["last_weight"]), int(cur.get("serialize", 0)),
                            int(cur.get("depth", 2)))


def current() -> dict:
    return {"backend": dict(_core._get_tuning()), "batch": dict(_core._get_batch_tuning())}


def apply(table: dict | None = None) -> dict:
    """Apply a tuning table (the committed one by default); returns it."""
    table = load() if table is None else table
    if table.get("backend"):
        set_backend(**{k: table["backend"][k] for k in BACKEND_KEYS if k in table["backend"]})
    if table.get("batch"):
        set_batch(**{k: table["batch"][k] for k in BATCH_KEYS if k in table["batch"]})
    return table
