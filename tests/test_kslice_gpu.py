"""GPU: k-sliced tiles of different tile classes of one contraction run
concurrently (on the class streams), each class in its own region of the
partial-sum buffer; results must match the oracle on every repetition.
Also a medium-size matvec chain, where the heavy contractions k-slice in
several classes at once."""
import numpy as np
import pytest

from oracle import oracle as O
from tensortoolkit_amd import capi, workloads as W
from tensortoolkit_amd.blocksparse import BlockSparseTensor, ContractionPlan, Index, IN, OUT
from tests.helpers import per_block_rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_two_sliced_classes_in_one_contraction(qlt, dtype):
    # sectors alternate 64 / 32 rows and columns, k = 2048 (128 steps): the
    # 64 x 64 outputs take tile class 0, the 32 x 32 outputs class 3, and
    # both are cut into k-slices
    S = 8
    qns = tuple((q,) for q in range(S))
    dm = tuple(64 if q % 2 == 0 else 32 for q in range(S))
    a = BlockSparseTensor.allowed_blocks([Index(qns, dm, IN), Index(qns, (2048,) * S, OUT)], (0,),
                                         dtype).fill_uniform(31)
    b = BlockSparseTensor.allowed_blocks([Index(qns, (2048,) * S, IN), Index(qns, dm, OUT)], (0,),
                                         dtype).fill_uniform(32)
    _say(f"case: two k-sliced tile classes, {dtype}")
    plan = ContractionPlan(a, b, [1], [0])
    assert plan.info.n_tiles > S            # sliced
    ref = O.OraclePlan(a, b, [1], [0]).execute()
    out = plan.output_tensor(dtype)
    da, db = capi.DeviceBuffer.from_numpy(a.data), capi.DeviceBuffer.from_numpy(b.data)
    dc = capi.DeviceBuffer(out.nbytes())
    for rep in range(5):
        plan.execute(dtype, capi.QLT_MODE_NATIVE, da.ptr, db.ptr, dc.ptr)
        capi.check(capi.lib().qlt_device_sync())
        out.data = dc.to_numpy(out.np_dtype, out.raw_size)
        assert per_block_rel_err(out, ref) <= TOL, rep


@pytest.mark.parametrize("mode", [capi.QLT_MODE_NATIVE, capi.QLT_MODE_SPLIT_BF16])
def test_medium_matvec_chain(qlt, mode):
    mv = W.small_matvec(D=768)
    _say(f"case: U(1) matvec chain D=768 f64 mode {mode}")
    x_ref = mv.psi
    x_dev = capi.DeviceBuffer.from_numpy(mv.psi.data)
    keep = [x_dev]
    for step, axes in enumerate(W.MATVEC_AXES):
        op = (mv.L, mv.W, mv.W, mv.R)[step]
        first, second = (op, x_ref) if step == 0 else (x_ref, op)
        ref = O.OraclePlan(first, second, list(axes[0]), list(axes[1]))
        out = ref.output_tensor(mv.dtype)
        out.data = ref.execute()
        plan = ContractionPlan(first, second, list(axes[0]), list(axes[1]))
        plan.prepare(mv.dtype, mode)
        dop = capi.DeviceBuffer.from_numpy(op.data)
        dc = capi.DeviceBuffer(out.nbytes())
        a_ptr, b_ptr = (dop.ptr, x_dev.ptr) if step == 0 else (x_dev.ptr, dop.ptr)
        plan.execute(mv.dtype, mode, a_ptr, b_ptr, dc.ptr)
        capi.check(capi.lib().qlt_device_sync())
        got = out.data.copy()
        got_t = plan.output_tensor(mv.dtype)
        got_t.data = dc.to_numpy(got_t.np_dtype, got_t.raw_size)
        err = per_block_rel_err(got_t, out.data)
        _say(f"  step {step}: tiles {plan.info.n_tiles}, rel err {err:.2e}")
        assert err <= TOL
        keep += [dop, dc]
        x_ref, x_dev = out, dc


@pytest.mark.parametrize("dtype,store", [("f64", "f64"), ("c128", "c128"), ("f64", "f32"), ("c128", "c64")])
def test_sliced_classes_alpha_beta(qlt, dtype, store):
    # diagonal blocks of 64, 32, 3 and 10 rows and columns at k = 2048: the
    # MFMA classes 0 and 3 and both row classes (n <= 4, n <= 16) are
    # k-sliced, so alpha and beta are applied by the slice-reduction kernels
    # (C read once per element); the unsliced tiles of the same classes run
    # at k = 16.  Every element type of the C ABI, single precision included.
    S = 8
    qns = tuple((q,) for q in range(S))
    dm = tuple((64, 32, 3, 10)[q % 4] for q in range(S))
    ks = tuple(2048 if q < 4 else 16 for q in range(S))
    a = BlockSparseTensor.allowed_blocks([Index(qns, dm, IN), Index(qns, ks, OUT)], (0,), dtype).fill_uniform(41)
    b = BlockSparseTensor.allowed_blocks([Index(qns, ks, IN), Index(qns, dm, OUT)], (0,), dtype).fill_uniform(42)
    npt = {"f64": np.float64, "c128": np.complex128, "f32": np.float32, "c64": np.complex64}[store]
    code = {"f64": capi.QLT_F64, "c128": capi.QLT_C128, "f32": capi.QLT_F32, "c64": capi.QLT_C64}[store]
    a.data = a.data.astype(npt).astype(a.np_dtype)
    b.data = b.data.astype(npt).astype(b.np_dtype)
    _say(f"case: k-sliced classes with alpha/beta, {dtype} stored as {store}")
    plan = ContractionPlan(a, b, [1], [0])
    assert plan.info.n_tiles > S            # sliced
    n = int(plan.info.c_raw_size)
    init = (np.arange(n) % 7 - 3.0) + (1j * (np.arange(n) % 5 - 2.0) if dtype == "c128" else 0.0)
    init = init.astype(npt)
    alpha, beta = (0.5 - 1.5j, 2.0 + 0.25j) if dtype == "c128" else (1.5, -0.75)
    da = capi.DeviceBuffer.from_numpy(a.data.astype(npt))
    db = capi.DeviceBuffer.from_numpy(b.data.astype(npt))
    exp = alpha * O.OraclePlan(a, b, [1], [0]).execute() + beta * init.astype(a.np_dtype)
    tol = 1e-12 if store in ("f64", "c128") else 2e-7
    for rep in range(3):
        dc = capi.DeviceBuffer.from_numpy(init)
        al = np.array([np.real(alpha), np.imag(alpha)])
        be = np.array([np.real(beta), np.imag(beta)])
        capi.check(capi.lib().qlt_contract_execute(plan.handle, code, capi.QLT_MODE_NATIVE, da.ptr, db.ptr, dc.ptr,
                                                   capi.f64(al), capi.f64(be), None))
        capi.check(capi.lib().qlt_device_sync())
        got = dc.to_numpy(npt, n).astype(a.np_dtype)
        assert np.linalg.norm(got - exp) <= tol * np.linalg.norm(exp), rep
