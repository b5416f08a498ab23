"""ctypes bindings for the Harvest oracle (test infrastructure only).

Two CPU libraries live here:
  * ``liboracle.so``  - the C restatement of WORLD Harvest (harvest_oracle.c),
    always built from this directory;
  * ``_ref/libworld_ref.so`` - WORLD itself compiled from the reference
    checkout (optional; absent where the reference checkout is absent).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg may
import this module, and only as the checker.  The product never loads it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_LIB = HERE / "liboracle.so"
_REF = HERE / "_ref" / "libworld_ref.so"
_DP = ctypes.POINTER(ctypes.c_double)


def build(ref: bool = True) -> None:
    """Compile liboracle.so (and the WORLD tap when the reference exists)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "all"], check=True)
    if ref and Path("/root/reference/third_party/world/src/harvest.cpp").exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_DP)


class _Params(ctypes.Structure):
    _fields_ = [("bands", ctypes.c_int), ("ratio", ctypes.c_int),
                ("fs_work", ctypes.c_double), ("y_length", ctypes.c_int),
                ("fft_size", ctypes.c_int), ("frames", ctypes.c_int),
                ("out_frames", ctypes.c_int), ("capacity", ctypes.c_int),
                ("slots", ctypes.c_int)]


@dataclass
class Params:
    bands: int
    ratio: int
    fs_work: float
    y_length: int
    fft_size: int
    frames: int
    out_frames: int
    capacity: int
    slots: int


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not _LIB.exists():
            build(ref=False)
        _lib = ctypes.CDLL(str(_LIB))
        _lib.ho_params_get.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double,
                                       ctypes.POINTER(_Params)]
    return _lib


def ref_available() -> bool:
    return _REF.exists()


def ref():
    global _ref
    if _ref is None:
        if not _REF.exists():
            raise FileNotFoundError("WORLD tap not built (oracle/_ref/libworld_ref.so)")
        _ref = ctypes.CDLL(str(_REF))
        _ref.ref_official_overlap.restype = ctypes.c_int
    return _ref


def params(length: int, fs: int, floor=71.0, ceil=800.0, period=5.0) -> Params:
    p = _Params()
    lib().ho_params_get(length, fs, floor, ceil, period, ctypes.byref(p))
    return Params(*(getattr(p, f[0]) for f in _Params._fields_))


def _cp(p: Params) -> _Params:
    return _Params(*(getattr(p, f[0]) for f in _Params._fields_))


def boundaries(p: Params, floor=71.0) -> np.ndarray:
    out = np.zeros(p.bands)
    cp = _cp(p)
    lib().ho_boundaries(ctypes.byref(cp), ctypes.c_double(floor), _ptr(out))
    return out


def _c(x):
    return np.ascontiguousarray(x, dtype=np.float64)


# ----------------------------------------------------------------- restatement
def decimate(x, fs, floor=71.0, ceil=800.0, remove_mean=True) -> np.ndarray:
    x = _c(x)
    p = params(len(x), fs, floor, ceil)
    y = np.zeros(p.y_length)
    lib().ho_decimate(_ptr(x), len(x), p.ratio, p.y_length, _ptr(y))
    if remove_mean:
        lib().ho_remove_mean(_ptr(y), p.y_length)
    return y


def filtered_band(y, p: Params, boundary: float) -> np.ndarray:
    y = _c(y)
    out = np.zeros(len(y))
    lib().ho_filtered_band(_ptr(y), len(y), p.fft_size, ctypes.c_double(p.fs_work),
                           ctypes.c_double(boundary), _ptr(out))
    return out


def band_candidates(y, p: Params, boundary, floor=71.0, ceil=800.0) -> np.ndarray:
    y = _c(y)
    out = np.zeros(p.frames)
    lib().ho_band_candidates(_ptr(y), len(y), p.fft_size, ctypes.c_double(p.fs_work),
                             ctypes.c_double(boundary), ctypes.c_double(floor),
                             ctypes.c_double(ceil), p.frames, _ptr(out))
    return out


def raw_candidates(y, p: Params, floor=71.0, ceil=800.0) -> np.ndarray:
    y = _c(y)
    b = boundaries(p, floor)
    raw = np.zeros((p.bands, p.frames))
    cp = _cp(p)
    lib().ho_raw_candidates(_ptr(y), ctypes.byref(cp), _ptr(b), ctypes.c_double(floor),
                            ctypes.c_double(ceil), _ptr(raw))
    return raw


def official(raw, p: Params):
    """(cand[F][slots] with official + overlapped candidates, count)."""
    raw = _c(raw)
    cand = np.zeros((p.frames, p.slots))
    cp = _cp(p)
    count = lib().ho_official(_ptr(raw), ctypes.byref(cp), _ptr(cand))
    lib().ho_overlap(ctypes.byref(cp), count, _ptr(cand))
    return cand, count


def refine(y, p: Params, cand, ncand, floor=71.0, ceil=800.0):
    y = _c(y)
    f0 = np.array(cand, dtype=np.float64, order="C", copy=True)
    score = np.zeros_like(f0)
    lib().ho_refine(_ptr(y), len(y), ctypes.c_double(p.fs_work), f0.shape[0],
                    f0.shape[1], ncand, ctypes.c_double(floor), ctypes.c_double(ceil),
                    _ptr(f0), _ptr(score))
    return f0, score


def refine_one(y, fs_work, t, f0, floor=71.0, ceil=800.0):
    y = _c(y)
    r, s = ctypes.c_double(), ctypes.c_double()
    lib().ho_refine_one(_ptr(y), len(y), ctypes.c_double(fs_work), ctypes.c_double(t),
                        ctypes.c_double(f0), ctypes.c_double(floor),
                        ctypes.c_double(ceil), ctypes.byref(r), ctypes.byref(s))
    return r.value, s.value


def prune(f0, score, ncand):
    f0 = np.array(f0, dtype=np.float64, order="C", copy=True)
    score = np.array(score, dtype=np.float64, order="C", copy=True)
    lib().ho_prune(f0.shape[0], f0.shape[1], ncand, _ptr(f0), _ptr(score))
    return f0, score


def base(f0, score, ncand) -> np.ndarray:
    f0, score = _c(f0), _c(score)
    out = np.zeros(f0.shape[0])
    lib().ho_base(_ptr(f0), _ptr(score), f0.shape[0], f0.shape[1], ncand, _ptr(out))
    return out


def fix_contour(f0, score, ncand) -> np.ndarray:
    f0, score = _c(f0), _c(score)
    out = np.zeros(f0.shape[0])
    lib().ho_fix_contour(_ptr(f0), _ptr(score), f0.shape[0], f0.shape[1], ncand,
                         _ptr(out))
    return out


def smooth(f0) -> np.ndarray:
    f0 = _c(f0)
    out = np.zeros(len(f0))
    lib().ho_smooth(_ptr(f0), len(f0), _ptr(out))
    return out


def harvest(x, fs, floor=71.0, ceil=800.0, period=5.0):
    x = _c(x)
    p = params(len(x), fs, floor, ceil, period)
    f0 = np.zeros(p.out_frames)
    t = np.zeros(p.out_frames)
    lib().ho_harvest(_ptr(x), len(x), fs, ctypes.c_double(floor), ctypes.c_double(ceil),
                     ctypes.c_double(period), _ptr(f0), _ptr(t))
    return f0, t


def stages(x, fs, floor=71.0, ceil=800.0):
    """Every intermediate of the restatement, for stage-level parity checks."""
    p = params(len(x), fs, floor, ceil)
    y = decimate(x, fs, floor, ceil)
    raw = raw_candidates(y, p, floor, ceil)
    cand, count = official(raw, p)
    ncand = 7 * count
    f0, score = refine(y, p, cand, ncand, floor, ceil)
    pf0, pscore = prune(f0, score, ncand)
    return dict(params=p, y=y, raw=raw, cand=cand, count=count, ncand=ncand,
                refined_f0=f0, refined_score=score, f0=pf0, score=pscore,
                best=base(pf0, pscore, ncand), contour=fix_contour(pf0, pscore, ncand))


# ---------------------------------------------------------------- WORLD (tap)
def ref_harvest(x, fs, floor=71.0, ceil=800.0, period=5.0):
    x = _c(x)
    p = params(len(x), fs, floor, ceil, period)
    f0 = np.zeros(p.out_frames)
    t = np.zeros(p.out_frames)
    ref().ref_harvest(_ptr(x), len(x), fs, ctypes.c_double(floor), ctypes.c_double(ceil),
                      ctypes.c_double(period), _ptr(f0), _ptr(t))
    return f0, t


def ref_waveform(x, fs, floor=71.0, ceil=800.0):
    x = _c(x)
    p = params(len(x), fs, floor, ceil)
    y = np.zeros(p.y_length)
    ref().ref_waveform(_ptr(x), len(x), fs, ctypes.c_double(floor), ctypes.c_double(ceil),
                       _ptr(y))
    return y


def ref_filtered_band(x, fs, band, floor=71.0, ceil=800.0):
    """WORLD's filtered band signal divided by n (exact: n is a power of two)."""
    x = _c(x)
    p = params(len(x), fs, floor, ceil)
    out = np.zeros(p.y_length)
    ref().ref_filtered_band(_ptr(x), len(x), fs, ctypes.c_double(floor),
                            ctypes.c_double(ceil), band, _ptr(out))
    return out / p.fft_size


def ref_raw_candidates(x, fs, floor=71.0, ceil=800.0):
    x = _c(x)
    p = params(len(x), fs, floor, ceil)
    raw = np.zeros((p.bands, p.frames))
    ref().ref_raw_candidates(_ptr(x), len(x), fs, ctypes.c_double(floor),
                             ctypes.c_double(ceil), _ptr(raw))
    return raw


def ref_official(raw, p: Params):
    raw = _c(raw)
    cand = np.zeros((p.frames, p.slots))
    count = ref().ref_official_overlap(_ptr(raw), p.bands, p.frames, p.slots, _ptr(cand))
    return cand, count


def ref_refine(y, p: Params, cand, ncand, floor=71.0, ceil=800.0):
    y = _c(y)
    f0 = np.array(cand, dtype=np.float64, order="C", copy=True)
    score = np.zeros_like(f0)
    ref().ref_refine(_ptr(y), len(y), ctypes.c_double(p.fs_work), f0.shape[0],
                     f0.shape[1], ncand, ctypes.c_double(floor), ctypes.c_double(ceil),
                     _ptr(f0), _ptr(score))
    return f0, score


def ref_prune(f0, score, ncand):
    f0 = np.array(f0, dtype=np.float64, order="C", copy=True)
    score = np.array(score, dtype=np.float64, order="C", copy=True)
    ref().ref_prune(f0.shape[0], f0.shape[1], ncand, _ptr(f0), _ptr(score))
    return f0, score


def ref_fix_contour(f0, score, ncand):
    f0 = np.array(f0, dtype=np.float64, order="C", copy=True)
    score = np.array(score, dtype=np.float64, order="C", copy=True)
    out = np.zeros(f0.shape[0])
    ref().ref_fix_contour(_ptr(f0), _ptr(score), f0.shape[0], f0.shape[1], ncand,
                          _ptr(out))
    return out


def ref_smooth(f0):
    f0 = _c(f0)
    out = np.zeros(len(f0))
    ref().ref_smooth(_ptr(f0), len(f0), _ptr(out))
    return out


def cpu_threads() -> int:
    return len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
