"""Autotuning of the grouped kernel's tile shapes and k-slicing on a workload.

Thin wrapper of ``qlt_autotune`` / ``qlt_tuning_get`` / ``qlt_tuning_set``
(no counterpart in the reference, whose per-task BLAS calls have no tiles).
A workload is a list of contractions with device operands; the tuner times it
under candidate settings (tile shape per size bucket (ceil(m/32), ceil(n/32)),
then the k-slicing workgroup target per tile class) and installs the fastest
as the library tuning, which every plan built afterwards uses.  A tuning can
be saved to JSON and loaded again in a later process.
"""
from __future__ import annotations

import ctypes as C
import json
from dataclasses import dataclass

import numpy as np

from . import capi
from .blocksparse import BlockSparseTensor, saved_axes


@dataclass
class TuneResult:
    tuning: dict
    ms_before: float      # workload time under the tuning in effect on entry
    ms_after: float       # workload time under the installed result
    trials: int           # candidate settings timed

    @property
    def speedup(self) -> float:
        return self.ms_before / self.ms_after if self.ms_after > 0 else 1.0


def tuning_to_dict(t: capi.Tuning) -> dict:
    return {"tile_class": [[int(t.tile_class[i][j]) for j in range(capi.QLT_TUNE_BUCKETS)]
                           for i in range(capi.QLT_TUNE_BUCKETS)],
            "slice_target": [int(t.slice_target[c]) for c in range(capi.QLT_TILE_CLASSES)]}


def tuning_from_dict(d: dict) -> capi.Tuning:
    t = capi.Tuning()
    tc = d["tile_class"]
    if len(tc) != capi.QLT_TUNE_BUCKETS or any(len(r) != capi.QLT_TUNE_BUCKETS for r in tc):
        raise ValueError(f"tile_class must be {capi.QLT_TUNE_BUCKETS} x {capi.QLT_TUNE_BUCKETS}")
    if len(d["slice_target"]) != capi.QLT_TILE_CLASSES:
        raise ValueError(f"slice_target must have {capi.QLT_TILE_CLASSES} entries")
    for i in range(capi.QLT_TUNE_BUCKETS):
        for j in range(capi.QLT_TUNE_BUCKETS):
            t.tile_class[i][j] = int(tc[i][j])
    for c in range(capi.QLT_TILE_CLASSES):
        t.slice_target[c] = int(d["slice_target"][c])
    return t


def get_tuning() -> dict:
    t = capi.Tuning()
    capi.check(capi.lib().qlt_tuning_get(C.byref(t)))
    return tuning_to_dict(t)


def set_tuning(d: dict | None) -> None:
    """Install a tuning (None: the built-in rules)."""
    if d is None:
        capi.check(capi.lib().qlt_tuning_set(None))
    else:
        t = tuning_from_dict(d)
        capi.check(capi.lib().qlt_tuning_set(C.byref(t)))


def overrides(d: dict) -> int:
    """Number of table entries and slice targets that differ from the built-in rules."""
    return sum(v >= 0 for row in d["tile_class"] for v in row) + sum(v != 0 for v in d["slice_target"])


def save(path: str, d: dict | None = None) -> None:
    with open(path, "w") as f:
        json.dump(get_tuning() if d is None else d, f)


def load(path: str) -> dict:
    with open(path) as f:
        d = json.load(f)
    set_tuning(d)
    return d


def autotune(contractions, dtype: str, mode: int, reps: int = 5, max_buckets: int = 24,
             stream=None) -> TuneResult:
    """Tune on a workload and install the result.

    ``contractions``: iterable of ``(a, b, a_ctrct, b_ctrct, a_ptr, b_ptr, c_ptr)``
    with BlockSparseTensor ``a``, ``b`` (saved axes ascending, as in Contract)
    and device pointers; every ``c_ptr`` buffer is overwritten."""
    capi.init()
    contractions = list(contractions)
    items = (capi.TuneItem * len(contractions))()
    keep = []      # the descriptors and axis arrays must outlive the call
    for i, (a, b, a_ax, b_ax, pa, pb, pc) in enumerate(contractions):
        if not isinstance(a, BlockSparseTensor) or not isinstance(b, BlockSparseTensor):
            raise TypeError("autotune: operands must be BlockSparseTensor")
        da, db = a.desc(), b.desc()
        axes = [np.ascontiguousarray(np.asarray(v, dtype=np.int32)) for v in
                (a_ax, b_ax, saved_axes(a.rank, a_ax), saved_axes(b.rank, b_ax))]
        keep += [a, b, da, db, axes]
        it = items[i]
        it.a, it.b = C.pointer(da), C.pointer(db)
        it.n_ctrct = len(axes[0])
        it.a_ctrct_axes, it.b_ctrct_axes = capi.i32(axes[0]), capi.i32(axes[1])
        it.n_a_saved, it.a_saved_axes = len(axes[2]), capi.i32(axes[2])
        it.n_b_saved, it.b_saved_axes = len(axes[3]), capi.i32(axes[3])
        it.a_data, it.b_data, it.c_data = pa, pb, pc
    best = capi.Tuning()
    before, after, trials = C.c_double(), C.c_double(), C.c_int32()
    dt = capi.QLT_C128 if dtype == "c128" else capi.QLT_F64
    capi.check(capi.lib().qlt_autotune(len(contractions), items, dt, mode, reps, max_buckets, stream,
                                       C.byref(best), C.byref(before), C.byref(after), C.byref(trials)))
    return TuneResult(tuning_to_dict(best), before.value, after.value, trials.value)
