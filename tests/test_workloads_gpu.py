"""GPU parity at the benchmark workloads' full sizes.

Whole matvec steps (L·ψ, ·W, ·W, ·R) of three chains: the D = 2048 U(1) FP64
chain, the lognormal-degeneracy U(1) FP64 chain (blocks 1 to 512 wide) and
the D = 4096 U(1)×U(1) complex128 chain.  Each of the four plans must equal
the oracle's plan field by field (up to 11 496 tasks).  In every arithmetic
mode, at the library's default dispatch, each step's output must lie within
1e-12 relative Frobenius error of the oracle's output block by block.  The
GPU chain feeds its own outputs forward; the oracle runs OpenBLAS per task.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tensortoolkit_amd import capi, workloads as W
from tensortoolkit_amd.blocksparse import ContractionPlan
from tests.helpers import per_block_rel_err
from tests.test_contract_gpu import assert_plan_equal

pytestmark = pytest.mark.gpu
TOL = 1e-12
MODES = {"native": capi.QLT_MODE_NATIVE, "split_bf16": capi.QLT_MODE_SPLIT_BF16,
         "split_i8": capi.QLT_MODE_SPLIT_I8}
WORKLOADS = {
    "u1_D2048_f64": lambda: W.config1(),
    "u1_lognormal_f64": lambda: W.config2(),
    "u1u1_D4096_c128": lambda: W.config3(),
}


def _say(msg):
    print(msg, flush=True)


def _threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


@pytest.fixture(scope="module")
def oracle_blas():
    O.use_openblas(_threads())
    yield
    O.use_naive_gemm()


_cache = {}


def workload(name):
    """The workload and the oracle's four step outputs, built once."""
    if name not in _cache:
        _cache.clear()
        mv = WORKLOADS[name]()
        steps, x = [], mv.psi
        for step, axes in enumerate(W.MATVEC_AXES):
            op = (mv.L, mv.W, mv.W, mv.R)[step]
            first, second = (op, x) if step == 0 else (x, op)
            ref = O.OraclePlan(first, second, list(axes[0]), list(axes[1]))
            out = ref.output_tensor(mv.dtype)
            out.data = ref.execute()
            steps.append((first, second, list(axes[0]), list(axes[1]), ref, out))
            x = out
        _cache[name] = (mv, steps)
    return _cache[name]


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("name", list(WORKLOADS))
def test_full_size_matvec_step(qlt, oracle_blas, name, mode):
    mv, steps = workload(name)
    np_dt = np.complex128 if mv.dtype == "c128" else np.float64
    dev = {k: capi.DeviceBuffer.from_numpy(getattr(mv, k).data) for k in ("L", "W", "R", "psi")}
    prev = dev["psi"]
    for step, (first, second, a_ax, b_ax, ref, ref_out) in enumerate(steps):
        _say(f"case: {name} mode={mode} step={step} tasks={len(ref.tasks())}")
        plan = ContractionPlan(first, second, a_ax, b_ax)
        assert_plan_equal(plan, ref)
        raw = int(plan.info.c_raw_size)
        # poison the output: beta == 0 must never read it
        dc = capi.DeviceBuffer.from_numpy(np.full(max(1, raw), np.nan, dtype=np_dt))
        a_ptr, b_ptr = (dev["L"].ptr, prev.ptr) if step == 0 else (prev.ptr, dev[("L", "W", "W", "R")[step]].ptr)
        plan.execute(mv.dtype, MODES[mode], a_ptr, b_ptr, dc.ptr)
        capi.check(capi.lib().qlt_device_sync())
        got = plan.output_tensor(mv.dtype)
        got.data = dc.to_numpy(np_dt, raw)
        err = per_block_rel_err(got, ref_out.data)
        _say(f"  max per-block rel err {err:.3e}")
        assert err <= TOL, (step, err)
        prev = dc
        plan.close()
