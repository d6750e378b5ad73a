"""Single-precision element types (QLTEN_Float / QLTEN_ComplexFloat,
framework/value_t.h:38-41) through the C ABI: float storage, FP64
arithmetic, one rounding per output element.  The oracle contracts the same
float-rounded inputs in FP64, so every output block must agree to float
rounding (relative error below 2^-23 per block, 1.2e-7); the split modes,
which emulate FP64, run these types on the native kernels (identical output);
the elementwise primitives (axpby, dot, conversion, conjugation, segment
scaling, the float fill) are checked against numpy."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from tensortoolkit_amd import capi
from tensortoolkit_amd.blocksparse import ContractionPlan
from tests.helpers import per_block_rel_err, random_case

pytestmark = pytest.mark.gpu


def _single(dtype):
    return (capi.QLT_C64, np.complex64) if dtype == "c128" else (capi.QLT_F32, np.float32)


def _contract_single(a, b, a_ax, b_ax, mode, alpha=1.0, beta=0.0, c_init=None):
    code, npt = _single(a.dtype)
    plan = ContractionPlan(a, b, a_ax, b_ax)
    raw = int(plan.info.c_raw_size)
    da = capi.DeviceBuffer.from_numpy(a.data.astype(npt))
    db = capi.DeviceBuffer.from_numpy(b.data.astype(npt))
    init = np.full(max(1, raw), np.nan, dtype=npt) if c_init is None else c_init.astype(npt)
    dc = capi.DeviceBuffer.from_numpy(init)
    al = np.array([np.real(alpha), np.imag(alpha)])
    be = np.array([np.real(beta), np.imag(beta)])
    capi.check(capi.lib().qlt_contract_execute(plan.handle, code, mode, da.ptr, db.ptr, dc.ptr, capi.f64(al),
                                               capi.f64(be), None))
    capi.check(capi.lib().qlt_device_sync())
    return plan, dc.to_numpy(npt, raw)


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_single_precision_contraction(qlt, seed, dtype):
    rng = np.random.default_rng(3100 + seed)
    a, b, a_ax, b_ax = random_case(rng, dtype=dtype, nq=1 + seed % 2, max_deg=20, min_deg=2, allow_scalar=False)
    _, npt = _single(dtype)
    # the float-rounded inputs, contracted in FP64 by the oracle
    a.data = a.data.astype(npt).astype(a.np_dtype)
    b.data = b.data.astype(npt).astype(b.np_dtype)
    print(f"case: single seed={seed} {dtype} ranks={a.rank},{b.rank} axes={a_ax},{b_ax}", flush=True)
    plan, got = _contract_single(a, b, a_ax, b_ax, capi.QLT_MODE_NATIVE)
    exp = O.OraclePlan(a, b, a_ax, b_ax).execute()
    c = plan.output_tensor(dtype)
    c.data = got.astype(c.np_dtype)
    assert per_block_rel_err(c, exp) <= 1.2e-7
    # the split modes run single-precision types natively: the same bits
    for mode in (capi.QLT_MODE_SPLIT_BF16, capi.QLT_MODE_SPLIT_I8):
        _, got2 = _contract_single(a, b, a_ax, b_ax, mode)
        np.testing.assert_array_equal(got2, got)


def test_single_precision_alpha_beta(qlt):
    rng = np.random.default_rng(77)
    a, b, a_ax, b_ax = random_case(rng, dtype="c128", max_deg=12, min_deg=2, allow_scalar=False)
    a.data = a.data.astype(np.complex64).astype(np.complex128)
    b.data = b.data.astype(np.complex64).astype(np.complex128)
    plan = ContractionPlan(a, b, a_ax, b_ax)
    n = int(plan.info.c_raw_size)
    init = ((np.arange(n) % 7 - 3.0) + 1j * (np.arange(n) % 5 - 2.0)).astype(np.complex64)
    alpha, beta = 0.5 - 1.5j, 2.0 + 0.25j
    _, got = _contract_single(a, b, a_ax, b_ax, capi.QLT_MODE_NATIVE, alpha, beta, init)
    exp = alpha * O.OraclePlan(a, b, a_ax, b_ax).execute() + beta * init.astype(np.complex128)
    assert np.linalg.norm(got - exp) <= 2e-7 * np.linalg.norm(exp)


def test_single_precision_primitives(qlt):
    lib = capi.lib()
    rng = np.random.default_rng(5)
    n = 10007
    x = (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(np.complex64)
    y = (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(np.complex64)
    dx, dy = capi.DeviceBuffer.from_numpy(x), capi.DeviceBuffer.from_numpy(y)
    a, b = np.array([0.5, -0.25]), np.array([2.0, 1.0])
    capi.check(lib.qlt_axpby(capi.QLT_C64, n, capi.f64(a), dx.ptr, capi.f64(b), dy.ptr, None))
    exp = (complex(*a) * x.astype(np.complex128) + complex(*b) * y.astype(np.complex128)).astype(np.complex64)
    np.testing.assert_allclose(dy.to_numpy(np.complex64), exp, rtol=0, atol=2e-7)
    out = (C.c_double * 2)()
    capi.check(lib.qlt_dot(capi.QLT_C64, n, dx.ptr, dx.ptr, out, None))
    assert abs(out[0] - np.vdot(x.astype(np.complex128), x.astype(np.complex128)).real) <= 1e-12 * out[0]
    # float -> complex64 -> complex128 and back down
    xr = rng.uniform(-1, 1, n).astype(np.float32)
    dxr = capi.DeviceBuffer.from_numpy(xr)
    dz = capi.DeviceBuffer(8 * n)
    capi.check(lib.qlt_convert(capi.QLT_F32, capi.QLT_C64, n, dxr.ptr, dz.ptr, None))
    np.testing.assert_array_equal(dz.to_numpy(np.complex64), xr.astype(np.complex64))
    dw = capi.DeviceBuffer(16 * n)
    capi.check(lib.qlt_convert(capi.QLT_C64, capi.QLT_C128, n, dz.ptr, dw.ptr, None))
    np.testing.assert_array_equal(dw.to_numpy(np.complex128), xr.astype(np.complex128))
    assert lib.qlt_convert(capi.QLT_C128, capi.QLT_F64, n, dw.ptr, dz.ptr, None) == 1   # complex -> real refused
    capi.check(lib.qlt_conjugate_c64(n, dx.ptr, None))
    capi.check(lib.qlt_device_sync())
    np.testing.assert_array_equal(dx.to_numpy(np.complex64), np.conj(x))
    # the float fill is the double fill rounded
    dd, df = capi.DeviceBuffer(8 * n), capi.DeviceBuffer(4 * n)
    capi.check(lib.qlt_fill_uniform(dd.ptr, n, 99, 7, -1.0, 1.0, None))
    capi.check(lib.qlt_fill_uniform_f32(df.ptr, n, 99, 7, -1.0, 1.0, None))
    capi.check(lib.qlt_device_sync())
    np.testing.assert_array_equal(df.to_numpy(np.float32), dd.to_numpy(np.float64).astype(np.float32))
