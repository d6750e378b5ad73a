"""GPU parity of the contraction path through the C ABI against the oracle.

Plan: bit-exact (every RawDataCtrctTask field, the output block map).
Values: relative Frobenius error of every output block <= 1e-12 against the
FP64 oracle (the required accuracy), for FP64 and complex128.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tensortoolkit_amd import capi, workloads as W
from tensortoolkit_amd.blocksparse import BlockSparseTensor, ContractionPlan, Index, IN, OUT
from tests.helpers import per_block_rel_err, random_case, rel_frobenius

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _say(msg):
    print(msg, flush=True)


def gpu_contract(a, b, a_ax, b_ax, mode=capi.QLT_MODE_NATIVE, a_saved=None, b_saved=None):
    plan = ContractionPlan(a, b, a_ax, b_ax, a_saved, b_saved)
    c = plan.output_tensor(a.dtype)
    raw = int(plan.info.c_raw_size)
    da = capi.DeviceBuffer.from_numpy(a.data)
    db = capi.DeviceBuffer.from_numpy(b.data)
    esz = 16 if a.dtype == "c128" else 8
    # poison the output: beta == 0 must never read it
    dc = capi.DeviceBuffer.from_numpy(np.full(max(1, raw), np.nan, dtype=a.np_dtype))
    plan.execute(a.dtype, mode, da.ptr, db.ptr, dc.ptr)
    capi.check(capi.lib().qlt_device_sync())
    c.data = dc.to_numpy(a.np_dtype, raw)
    return plan, c


def assert_plan_equal(plan, ref):
    got, exp = plan.tasks(), ref.tasks()
    assert len(got) == len(exp)
    for name in got.dtype.names:
        np.testing.assert_array_equal(got[name], exp[name], err_msg=name)
    for g, e in zip(plan.c_blocks(), ref.c_blocks()):
        np.testing.assert_array_equal(g, e)
    assert plan.info.c_raw_size == ref.c_raw_size


def test_plan_and_values_rank3_u1(qlt):
    _say("case: rank-3 U(1) pair, f64")
    a, b, ax = W.config0()
    plan, c = gpu_contract(a, b, *ax)
    ref = O.OraclePlan(a, b, *ax)
    assert_plan_equal(plan, ref)
    assert plan.info.n_tasks == 32
    assert per_block_rel_err(c, ref.execute()) <= TOL


@pytest.mark.parametrize("seed", range(16))
@pytest.mark.parametrize("nq", [1, 2])
def test_plan_bit_exact_random(qlt, seed, nq):
    rng = np.random.default_rng(5000 + seed + 97 * nq)
    a, b, a_ax, b_ax = random_case(rng, nq=nq, max_rank=5, max_sectors=5)
    _say(f"case: seed={seed} nq={nq} ranks={a.rank},{b.rank} axes={a_ax},{b_ax} blocks={a.n_blks},{b.n_blks}")
    plan = ContractionPlan(a, b, a_ax, b_ax)
    assert_plan_equal(plan, O.OraclePlan(a, b, a_ax, b_ax))


@pytest.mark.parametrize("seed", range(12))
@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_values_random(qlt, seed, dtype):
    rng = np.random.default_rng(7000 + seed)
    a, b, a_ax, b_ax = random_case(rng, dtype=dtype, nq=1 + seed % 2, max_deg=9)
    _say(f"case: seed={seed} {dtype} ranks={a.rank},{b.rank} axes={a_ax},{b_ax}")
    plan, c = gpu_contract(a, b, a_ax, b_ax)
    ref = O.OraclePlan(a, b, a_ax, b_ax)
    exp = ref.execute()
    if plan.info.c_is_scalar:
        assert abs(c.data[0] - exp[0]) <= TOL * max(1.0, abs(exp[0]))
    else:
        assert per_block_rel_err(c, exp) <= TOL


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_values_random_wide_blocks(qlt, seed, dtype):
    """Degeneracies up to 40, so many output blocks have n > 16 and take
    the MFMA tile class (narrower ones take the row class)."""
    rng = np.random.default_rng(7500 + seed)
    a, b, a_ax, b_ax = random_case(rng, dtype=dtype, nq=1 + seed % 2, max_rank=3, max_sectors=3,
                                   max_deg=40, min_deg=8, allow_scalar=False)
    _say(f"case: wide seed={seed} {dtype} ranks={a.rank},{b.rank} axes={a_ax},{b_ax}")
    plan, c = gpu_contract(a, b, a_ax, b_ax)
    ref = O.OraclePlan(a, b, a_ax, b_ax)
    assert_plan_equal(plan, ref)
    assert per_block_rel_err(c, ref.execute()) <= TOL


@pytest.mark.parametrize("scale", [1e-8, 1e8])
@pytest.mark.parametrize("mode", [capi.QLT_MODE_NATIVE, capi.QLT_MODE_SPLIT_BF16, capi.QLT_MODE_SPLIT_I8])
def test_complex_parts_of_very_different_size(qlt, scale, mode):
    """Complex128 blocks whose imaginary parts are 1e-8 (or 1e8) times their
    real parts, in A and in B: the native kernel's three-product form builds
    im = (ar+ai)(br+bi) - ar br - ai bi, whose rounding scales with the larger
    part, and the split modes share one exponent per row over both parts;
    every output block stays within 1e-12 of the oracle in Frobenius norm."""
    rng = np.random.default_rng(9100)
    a, b, a_ax, b_ax = random_case(rng, dtype="c128", nq=1, max_rank=3, max_sectors=3, max_deg=40, min_deg=8,
                                   allow_scalar=False)
    a.data = a.data.real + 1j * scale * a.data.imag
    b.data = b.data.real + 1j * scale * b.data.imag
    _say(f"case: unbalanced complex parts scale={scale} mode={mode}")
    old = capi.lib().qlt_get_split_min_mean_k()
    capi.check(capi.lib().qlt_set_split_min_mean_k(0.0))
    try:
        plan, c = gpu_contract(a, b, a_ax, b_ax, mode)
    finally:
        capi.check(capi.lib().qlt_set_split_min_mean_k(old))
    assert per_block_rel_err(c, O.OraclePlan(a, b, a_ax, b_ax).execute()) <= TOL


@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_large_blocks_and_ragged_edges(qlt, dtype):
    """Blocks bigger than one 64x64 tile with ragged edges and long k."""
    ix_l = Index(((0,), (1,)), (70, 131), IN)
    ix_r = Index(((0,), (1,)), (129, 67), OUT)
    ix_p = Index(((0,),), (3,), IN)
    a = BlockSparseTensor.allowed_blocks([ix_l, ix_p, ix_r], (0,), dtype).fill_uniform(11)
    b = BlockSparseTensor.allowed_blocks([ix_r.inverse(), ix_p, ix_l.inverse()], (0,), dtype).fill_uniform(12)
    _say(f"case: ragged {dtype}")
    for axes in (([2], [0]), ([0], [2]), ([2, 1], [0, 1]), ([1, 2], [1, 0])):
        plan, c = gpu_contract(a, b, *axes)
        ref = O.OraclePlan(a, b, *axes)
        assert_plan_equal(plan, ref)
        assert per_block_rel_err(c, ref.execute()) <= TOL, axes


def test_beta_accumulates(qlt):
    a, b, ax = W.config0()
    plan = ContractionPlan(a, b, *ax)
    da = capi.DeviceBuffer.from_numpy(a.data)
    db = capi.DeviceBuffer.from_numpy(b.data)
    init = np.linspace(-1, 1, plan.info.c_raw_size)
    dc = capi.DeviceBuffer.from_numpy(init)
    plan.execute("f64", capi.QLT_MODE_NATIVE, da.ptr, db.ptr, dc.ptr, alpha=2.0, beta=-0.5)
    got = dc.to_numpy(np.float64)
    exp = 2.0 * O.OraclePlan(a, b, *ax).execute() - 0.5 * init
    assert rel_frobenius(got, exp) <= TOL


def test_complex_alpha_beta(qlt):
    rng = np.random.default_rng(3)
    a, b, a_ax, b_ax = random_case(rng, dtype="c128", allow_scalar=False, max_deg=7)
    plan = ContractionPlan(a, b, a_ax, b_ax)
    n = plan.info.c_raw_size
    init = (np.arange(n) % 7 - 3.0) + 1j * (np.arange(n) % 5 - 2.0)
    da, db = capi.DeviceBuffer.from_numpy(a.data), capi.DeviceBuffer.from_numpy(b.data)
    dc = capi.DeviceBuffer.from_numpy(init.astype(np.complex128))
    alpha, beta = 0.5 - 1.5j, 2.0 + 0.25j
    plan.execute("c128", capi.QLT_MODE_NATIVE, da.ptr, db.ptr, dc.ptr, alpha=alpha, beta=beta)
    got = dc.to_numpy(np.complex128)
    exp = alpha * O.OraclePlan(a, b, a_ax, b_ax).execute() + beta * init
    assert rel_frobenius(got, exp) <= TOL


def test_no_matching_blocks_gives_empty_plan(qlt):
    ix = Index(((0,), (1,)), (2, 2), OUT)
    a = BlockSparseTensor.allowed_blocks([ix, ix], (0,)).fill_uniform(1)      # only (0,0)
    b = BlockSparseTensor.allowed_blocks([ix.inverse(), ix.inverse()], (-2,)).fill_uniform(2)  # only (1,1)
    plan = ContractionPlan(a, b, [1], [0])
    assert plan.info.n_tasks == 0 and plan.info.c_raw_size == 0
    plan.execute("f64", capi.QLT_MODE_NATIVE, None, None, None)   # a no-op


def test_device_fill_matches_host(qlt):
    n = 10007
    buf = capi.DeviceBuffer(8 * n)
    capi.check(capi.lib().qlt_fill_uniform(buf.ptr, n, 99, 3, -1.0, 1.0, None))
    capi.check(capi.lib().qlt_device_sync())
    np.testing.assert_array_equal(buf.to_numpy(np.float64), O.fill_uniform(n, 99, 3))


@pytest.mark.parametrize("u1u1,dtype", [(False, "f64"), (True, "c128")])
def test_small_matvec_chain(qlt, u1u1, dtype):
    mv = W.small_matvec(D=64 if u1u1 else 96, dtype=dtype, u1u1=u1u1)
    _say(f"case: small matvec {mv.name} {dtype}")
    x = mv.psi
    x_ref = mv.psi
    for step, (ops, axes) in enumerate(zip((mv.L, mv.W, mv.W, mv.R), W.MATVEC_AXES)):
        first, second = (ops, x) if step == 0 else (x, ops)
        first_ref, second_ref = (ops, x_ref) if step == 0 else (x_ref, ops)
        plan, c = gpu_contract(first, second, list(axes[0]), list(axes[1]))
        ref = O.OraclePlan(first_ref, second_ref, list(axes[0]), list(axes[1]))
        assert_plan_equal(plan, ref)
        c_ref = ref.output_tensor(dtype)
        c_ref.data = ref.execute()
        assert per_block_rel_err(c, c_ref.data) <= TOL, step
        x, x_ref = c, c_ref
    assert [ix.dir for ix in x.indexes] == [IN, IN, IN, OUT]


def test_kernel_probe_names_and_counts_dominant_kernel(qlt):
    """The probe brackets one kernel and reports its algorithmic flops: the
    rank-3 U(1) pair has 32 tasks of (16,16,16), all on the n <= 16 row kernel,
    2 * 32 * 16**3 = 262144 flops (the 2.62e5 of the workload definition)."""
    import ctypes as C
    _say("case: kernel probe on the rank-3 U(1) pair")
    hip = C.CDLL("libamdhip64.so")
    hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipEventSynchronize.argtypes = [C.c_void_p]
    hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
    hip.hipEventDestroy.argtypes = [C.c_void_p]
    e0, e1 = C.c_void_p(), C.c_void_p()
    assert hip.hipEventCreate(C.byref(e0)) == 0 and hip.hipEventCreate(C.byref(e1)) == 0
    lib = capi.lib()
    a, b, ax = W.config0()
    capi.check(lib.qlt_kernel_probe_set(e0, e1))
    try:
        plan, c = gpu_contract(a, b, *ax)
        name = C.create_string_buffer(96)
        fl, by = C.c_double(), C.c_double()
        capi.check(lib.qlt_kernel_probe_last(name, 96, C.byref(fl), C.byref(by)))
    finally:
        capi.check(lib.qlt_kernel_probe_set(None, None))
    assert hip.hipEventSynchronize(e1) == 0
    ms = C.c_float()
    assert hip.hipEventElapsedTime(C.byref(ms), e0, e1) == 0
    assert ms.value > 0.0
    assert name.value.decode() == "k_rows_f64<n<=16> f64"
    assert fl.value == 262144.0
    t = plan.tasks()
    elems = 32 * 256 + 256 * len(set(t["a_blk_idx"])) + 256 * len(set(t["b_blk_idx"]))
    assert by.value == 8.0 * elems
    # disarmed: a later call records nothing
    assert lib.qlt_kernel_probe_last(name, 96, C.byref(fl), C.byref(by)) != 0
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
