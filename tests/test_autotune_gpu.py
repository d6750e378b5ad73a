"""GPU: the tuning table and slice targets reshape the device work list
without touching the plan or the values, and the autotuner's result keeps
the contraction within 1e-12 of the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from tensortoolkit_amd import autotune as T
from tensortoolkit_amd import capi, workloads as W
from tensortoolkit_amd.blocksparse import BlockSparseTensor, ContractionPlan, Index, IN, OUT
from tests.helpers import per_block_rel_err
from tests.test_contract_gpu import assert_plan_equal, gpu_contract

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _say(msg):
    print(msg, flush=True)


@pytest.fixture
def builtin_tuning(qlt):
    T.set_tuning(None)
    yield
    T.set_tuning(None)


def _diag_pair(d_m, d_k, d_n, S, dtype="f64", seed=1):
    qns = tuple((q,) for q in range(S))
    a = BlockSparseTensor.allowed_blocks([Index(qns, (d_m,) * S, IN), Index(qns, (d_k,) * S, OUT)], (0,),
                                         dtype).fill_uniform(seed)
    b = BlockSparseTensor.allowed_blocks([Index(qns, (d_k,) * S, IN), Index(qns, (d_n,) * S, OUT)], (0,),
                                         dtype).fill_uniform(seed + 1)
    return a, b


@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_bucket_table_sets_the_tile_shape(builtin_tuning, dtype):
    # 96 x 96 outputs, k = 96 (6 steps, too short to be k-sliced): the
    # built-in rule takes 32 x 32 tiles (9 per block), bucket (3, 3) forced
    # to 64 x 64 gives 4 per block.
    _say(f"case: tuning table, 8 blocks 96x96x96 {dtype}")
    a, b = _diag_pair(96, 96, 96, 8, dtype)
    ref = O.OraclePlan(a, b, [1], [0])
    plan0, c0 = gpu_contract(a, b, [1], [0])
    assert plan0.info.n_tiles == 8 * 9
    d = T.get_tuning()
    d["tile_class"][2][2] = 0
    T.set_tuning(d)
    plan1, c1 = gpu_contract(a, b, [1], [0])
    assert plan1.info.n_tiles == 8 * 4
    assert_plan_equal(plan1, ref)
    exp = ref.execute()
    assert per_block_rel_err(c0, exp) <= TOL
    assert per_block_rel_err(c1, exp) <= TOL


def test_other_buckets_keep_the_builtin_rule(builtin_tuning):
    _say("case: tuning table entry of another bucket, 8 blocks 96x96x96")
    a, b = _diag_pair(96, 96, 96, 8)
    d = T.get_tuning()
    d["tile_class"][2][3] = 0      # bucket (3, 4): no block here
    T.set_tuning(d)
    plan, _ = gpu_contract(a, b, [1], [0])
    assert plan.info.n_tiles == 8 * 9


@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_unsliced_target_keeps_one_tile_per_output_tile(builtin_tuning, dtype):
    # one 64 x 64 output per block with k = 2048 (128 steps), which the
    # built-in rule covers with two 32 x 64 tiles (class 1): the built-in
    # target cuts each tile into k-slices; -1 keeps them whole
    _say(f"case: slice target, 4 blocks 64x2048x64 {dtype}")
    a, b = _diag_pair(64, 2048, 64, 4, dtype)
    ref = O.OraclePlan(a, b, [1], [0])
    exp = ref.execute()
    plan0, c0 = gpu_contract(a, b, [1], [0])
    assert plan0.info.n_tiles > 8
    d = T.get_tuning()
    d["slice_target"][1] = -1
    T.set_tuning(d)
    plan1, c1 = gpu_contract(a, b, [1], [0])
    assert plan1.info.n_tiles == 8
    assert_plan_equal(plan1, ref)
    assert per_block_rel_err(c0, exp) <= TOL
    assert per_block_rel_err(c1, exp) <= TOL


@pytest.mark.parametrize("mode", [capi.QLT_MODE_NATIVE, capi.QLT_MODE_SPLIT_BF16, capi.QLT_MODE_SPLIT_I8])
def test_autotune_small_matvec(builtin_tuning, mode):
    _say(f"case: autotune small U(1) matvec D=96 mode {mode}")
    mv = W.small_matvec(D=96)
    x = mv.psi
    items, bufs, refs = [], [], []
    for step, axes in enumerate(W.MATVEC_AXES):
        op = (mv.L, mv.W, mv.W, mv.R)[step]
        first, second = (op, x) if step == 0 else (x, op)
        ref = O.OraclePlan(first, second, list(axes[0]), list(axes[1]))
        out = ref.output_tensor(mv.dtype)
        da, db = capi.DeviceBuffer.from_numpy(first.data), capi.DeviceBuffer.from_numpy(second.data)
        dc = capi.DeviceBuffer(8 * max(1, out.raw_size))
        bufs += [da, db, dc]
        items.append((first, second, list(axes[0]), list(axes[1]), da.ptr, db.ptr, dc.ptr))
        refs.append((first, second, list(axes), ref))
        out.data = ref.execute()
        x = out
    res = T.autotune(items, "f64", mode, reps=2, max_buckets=4)
    _say(f"autotune: {res.trials} trials, {res.ms_before:.3f} -> {res.ms_after:.3f} ms, "
         f"{T.overrides(res.tuning)} overrides")
    assert res.trials >= 2 and res.ms_before > 0 and res.ms_after > 0
    assert T.get_tuning() == res.tuning
    # the tuned library still reproduces every contraction of the chain
    for first, second, axes, ref in refs:
        plan, c = gpu_contract(first, second, axes[0], axes[1], mode=mode)
        assert_plan_equal(plan, ref)
        assert per_block_rel_err(c, ref.execute()) <= TOL
