"""Writes the plan fixtures in tests/golden/ (run from the repository root:
``python tests/golden/make_golden.py``).

Each fixture pins the reference plan (every RawDataCtrctTask field and the
output block map, data_blk_operations.h:411-578) of one contraction case whose
inputs are regenerated deterministically by tests/golden/cases.py.  The plans
come from the CPU oracle and are cross-checked here against the brute-force
pair enumeration before they are written.  The U(1)xU(1) case also stores the
output values (oracle FP64) as a regression fixture.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from tests.golden import cases  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden")


def dump(name, a, b, a_ax, b_ax, with_values=False):
    p = O.OraclePlan(a, b, a_ax, b_ax)
    t = p.tasks()
    bf = O.brute_force_plan(a, b, a_ax, b_ax)
    assert len(bf) == len(t), name
    ints = np.stack([t[f] for f in t.dtype.names[:9]], axis=1).astype(np.int64)
    flts = np.stack([t["f_ex_sign"], t["beta"]], axis=1)
    idx, coords, dims, offs = p.c_blocks()
    arrays = dict(task_ints=ints, task_f_beta=flts, c_blk_idx=idx, c_coords=coords, c_dims=dims,
                  c_offsets=offs, c_raw_size=np.array([p.c_raw_size]))
    if with_values:
        arrays["c_values"] = p.execute()
    np.savez_compressed(os.path.join(OUT, f"plan_{name}.npz"), **arrays)
    print(f"{name}: {len(t)} tasks, {len(idx)} output blocks")


def main():
    for name, (a, b, a_ax, b_ax, vals) in cases.all_cases().items():
        dump(name, a, b, a_ax, b_ax, vals)


if __name__ == "__main__":
    main()
