"""GPU parity of the contiguous-axes path (ContractContiguousAxes,
contract_contiguous_axes.h:305-353, 849-868): contracted axes
(start + i) % rank and saved axes cyclic from the end of the run, passed to
the device plan builder as explicit saved axes.  The plan is compared field
by field with the oracle's plan under the same explicit axes, the values
with the oracle per output block (and the oracle itself with the dense
contraction in tests/test_oracle.py).  The C++ forms of this API, all four
(ASide, BSide) template arguments and the legacy bool forms, are checked in
tests/cpp/test_host_api.cc."""
import numpy as np
import pytest

from oracle import oracle as O
from tensortoolkit_amd import capi
from tests.contiguous_cases import SHAPES, contiguous_case, dense_reference
from tests.helpers import per_block_rel_err
from tests.test_contract_gpu import assert_plan_equal, gpu_contract

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.mark.parametrize("shape", SHAPES, ids=[",".join(map(str, s)) for s in SHAPES])
@pytest.mark.parametrize("nq", [1, 2])
def test_contiguous_plan_bit_exact(qlt, shape, nq):
    rng = np.random.default_rng(4100 + 13 * SHAPES.index(shape) + nq)
    a, b, axes = contiguous_case(rng, shape, nq=nq, max_deg=5)
    print(f"case: contiguous plan shape={shape} nq={nq} axes={axes}", flush=True)
    from tensortoolkit_amd.blocksparse import ContractionPlan
    plan = ContractionPlan(a, b, *axes)
    assert_plan_equal(plan, O.OraclePlan(a, b, *axes))
    assert [ix for ix in plan.c_indexes] == [a.indexes[i] for i in axes[2]] + [b.indexes[i] for i in axes[3]]


@pytest.mark.parametrize("mode", [capi.QLT_MODE_NATIVE, capi.QLT_MODE_SPLIT_BF16, capi.QLT_MODE_SPLIT_I8],
                         ids=["native", "bf16", "i8"])
@pytest.mark.parametrize("dtype", ["f64", "c128"])
@pytest.mark.parametrize("shape", SHAPES, ids=[",".join(map(str, s)) for s in SHAPES])
def test_contiguous_values(qlt, shape, dtype, mode):
    lib = capi.lib()
    old = lib.qlt_get_split_min_mean_k()
    rng = np.random.default_rng(4700 + 7 * SHAPES.index(shape) + (dtype == "c128"))
    # blocks wide enough that the split kernels get MFMA tiles, dense
    # references of a few million elements at most
    r = max(shape[0], shape[1])
    a, b, axes = contiguous_case(rng, shape, dtype=dtype, nq=1, max_deg={3: 14, 4: 8, 5: 5}[r],
                                 min_deg=4 if r == 3 else 2)
    print(f"case: contiguous values shape={shape} {dtype} mode={mode} axes={axes}", flush=True)
    capi.check(lib.qlt_set_split_min_mean_k(0.0))
    try:
        plan, c = gpu_contract(a, b, axes[0], axes[1], mode=mode, a_saved=axes[2], b_saved=axes[3])
    finally:
        capi.check(lib.qlt_set_split_min_mean_k(old))
    ref = O.OraclePlan(a, b, *axes)
    exp = ref.execute()
    if plan.info.c_is_scalar:
        assert abs(c.data[0] - exp[0]) <= TOL * max(1.0, abs(exp[0]))
        return
    assert per_block_rel_err(c, exp) <= TOL
    d = dense_reference(a, b, axes)
    np.testing.assert_allclose(c.to_dense(), d, rtol=0, atol=1e-12 * max(1.0, np.abs(d).max()))
