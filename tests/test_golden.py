"""Golden plan fixtures (tests/golden/plan_*.npz, written by
tests/golden/make_golden.py from the inputs in tests/golden/cases.py).

CPU: the oracle reproduces every stored plan field; the stored output values
of the U(1)xU(1) case and the cost known-answer case agree with a dense
np.tensordot of the densified inputs.
GPU: the device plan builder reproduces every stored plan bit-exactly, and the
native contraction reproduces the stored values to 1e-12.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tensortoolkit_amd import capi
from tests.golden import cases

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = cases.all_cases()
NAMES = sorted(CASES)


def load(name):
    with np.load(os.path.join(HERE, f"plan_{name}.npz")) as z:
        return {k: z[k] for k in z.files}


def plan_fields(tasks, c_blocks, c_raw_size):
    ints = np.stack([tasks[f] for f in tasks.dtype.names[:9]], axis=1).astype(np.int64)
    flts = np.stack([tasks["f_ex_sign"], tasks["beta"]], axis=1)
    idx, coords, dims, offs = c_blocks
    return dict(task_ints=ints, task_f_beta=flts, c_blk_idx=idx, c_coords=coords, c_dims=dims,
                c_offsets=offs, c_raw_size=np.array([c_raw_size]))


def assert_fields_equal(got, exp, name):
    for key in ("task_ints", "task_f_beta", "c_blk_idx", "c_coords", "c_dims", "c_offsets", "c_raw_size"):
        np.testing.assert_array_equal(np.asarray(got[key]).reshape(np.asarray(exp[key]).shape), exp[key],
                                      err_msg=f"{name}:{key}")


def test_fixtures_present():
    for name in NAMES:
        assert os.path.exists(os.path.join(HERE, f"plan_{name}.npz")), name


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden_plan(name):
    a, b, a_ax, b_ax, _ = CASES[name]
    p = O.OraclePlan(a, b, a_ax, b_ax)
    assert_fields_equal(plan_fields(p.tasks(), p.c_blocks(), p.c_raw_size), load(name), name)


@pytest.mark.parametrize("name", [n for n in NAMES if CASES[n][4]])
def test_golden_values_match_dense(name):
    a, b, a_ax, b_ax, _ = CASES[name]
    g = load(name)
    p = O.OraclePlan(a, b, a_ax, b_ax)
    c = p.output_tensor("f64")
    c.data = g["c_values"]
    dense = O.dense_contract(a, b, a_ax, b_ax)
    np.testing.assert_allclose(c.to_dense(), dense, rtol=0, atol=1e-13 * max(1.0, np.abs(dense).max()))


def test_u1u1_multiblock_shape():
    # lhs and rhs carry 3 sectors on each of 4 legs with divergence 0; the
    # output (lhs legs 0,1 + rhs legs 2,3) keeps divergence 0.
    g = load("u1u1_multiblock")
    assert g["task_ints"].shape[0] == 45 and g["c_blk_idx"].shape[0] == 19


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_device_plan_matches_golden(qlt, name):
    from tensortoolkit_amd.blocksparse import ContractionPlan
    print(f"case: golden {name}", flush=True)
    a, b, a_ax, b_ax, with_values = CASES[name]
    g = load(name)
    plan = ContractionPlan(a, b, a_ax, b_ax)
    assert_fields_equal(plan_fields(plan.tasks(), plan.c_blocks(), int(plan.info.c_raw_size)), g, name)
    if with_values:
        raw = int(plan.info.c_raw_size)
        da = capi.DeviceBuffer.from_numpy(a.data)
        db = capi.DeviceBuffer.from_numpy(b.data)
        dc = capi.DeviceBuffer.from_numpy(np.full(raw, np.nan))
        plan.execute("f64", capi.QLT_MODE_NATIVE, da.ptr, db.ptr, dc.ptr)
        capi.check(capi.lib().qlt_device_sync())
        got = dc.to_numpy(np.float64, raw)
        err = np.linalg.norm(got - g["c_values"]) / np.linalg.norm(g["c_values"])
        assert err <= 1e-12
