"""GPU parity of the split (FP64-emulating) paths, BF16 slices and int8
digits, against the FP64 oracle.

Every output block must be within 1e-12 relative Frobenius error of the
oracle (the required accuracy), and the plan is the same plan as the native
path.  The split kernel is forced (dispatch threshold 0) so that small test
cases exercise it rather than the native fallback for low-k contractions.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tensortoolkit_amd import capi, workloads as W
from tensortoolkit_amd.blocksparse import BlockSparseTensor, ContractionPlan, Index, IN, OUT
from tests.helpers import per_block_rel_err, random_case, rel_frobenius
from tests.test_contract_gpu import gpu_contract

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(params=[capi.QLT_MODE_SPLIT_BF16, capi.QLT_MODE_SPLIT_I8], ids=["bf16", "i8"])
def split_mode(request):
    return request.param


@pytest.fixture
def forced_split(qlt):
    lib = capi.lib()
    old = lib.qlt_get_split_min_mean_k()
    capi.check(lib.qlt_set_split_min_mean_k(0.0))
    yield
    capi.check(lib.qlt_set_split_min_mean_k(old))


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("seed", range(10))
@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_split_random(forced_split, split_mode, seed, dtype):
    rng = np.random.default_rng(9000 + seed)
    a, b, a_ax, b_ax = random_case(rng, dtype=dtype, nq=1 + seed % 2, max_rank=3, max_sectors=3,
                                   max_deg=40, min_deg=8, allow_scalar=False)
    _say(f"case: split mode={split_mode} seed={seed} {dtype} ranks={a.rank},{b.rank} axes={a_ax},{b_ax}")
    plan, c = gpu_contract(a, b, a_ax, b_ax, mode=split_mode)
    ref = O.OraclePlan(a, b, a_ax, b_ax)
    err = per_block_rel_err(c, ref.execute())
    assert err <= TOL, err


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_split_random_rank4(forced_split, split_mode, seed, dtype):
    """Ranks up to 4 with contracted axes in random order: the gathers of
    permuted, non-plain blocks feed the split pack."""
    rng = np.random.default_rng(9500 + seed)
    while True:   # draw until the case has output blocks wide enough for the split kernels
        a, b, a_ax, b_ax = random_case(rng, dtype=dtype, nq=1 + seed % 2, max_rank=4, max_sectors=3,
                                       max_deg=20, min_deg=6, allow_scalar=False)
        t = O.OraclePlan(a, b, a_ax, b_ax).tasks()
        if len(t) and (t["n"] > 16).any() and (t["m"] > 16).any():
            break
    _say(f"case: split rank4 mode={split_mode} seed={seed} {dtype} ranks={a.rank},{b.rank} axes={a_ax},{b_ax}")
    plan, c = gpu_contract(a, b, a_ax, b_ax, mode=split_mode)
    ref = O.OraclePlan(a, b, a_ax, b_ax)
    err = per_block_rel_err(c, ref.execute())
    assert err <= TOL, err


@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_split_long_k_many_flushes(forced_split, split_mode, dtype):
    """k = 3 * 200 = 600 spans many FP32 flush intervals; ragged m, n."""
    ix_l = Index(((0,), (1,)), (70, 131), IN)
    ix_r = Index(((0,), (1,)), (200, 67), OUT)
    ix_p = Index(((0,),), (3,), IN)
    a = BlockSparseTensor.allowed_blocks([ix_l, ix_p, ix_r], (0,), dtype).fill_uniform(21)
    b = BlockSparseTensor.allowed_blocks([ix_r.inverse(), ix_p, ix_l.inverse()], (0,), dtype).fill_uniform(22)
    for axes in (([2, 1], [0, 1]), ([0], [2])):
        _say(f"case: split long-k {dtype} axes={axes}")
        plan, c = gpu_contract(a, b, *axes, mode=split_mode)
        ref = O.OraclePlan(a, b, *axes)
        err = per_block_rel_err(c, ref.execute())
        assert err <= TOL, (axes, err)


@pytest.mark.parametrize("k", [40, 64, 100, 256, 300, 512, 513, 1024, 1025])
@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_split_pack_widths(forced_split, split_mode, dtype, k):
    """Every single-read pack width (k up to 64, 128, 256, 512, 1024) at and
    around its limits, and the two-pass pack beyond 1024; ragged rows,
    transposed B."""
    ix_m = Index(((0,),), (77,), IN)
    ix_k = Index(((0,),), (k,), OUT)
    ix_n = Index(((0,),), (45,), OUT)
    a = BlockSparseTensor.allowed_blocks([ix_m, ix_k], (0,), dtype).fill_uniform(61)
    b = BlockSparseTensor.allowed_blocks([ix_n.inverse(), ix_k.inverse()], (0,), dtype).fill_uniform(62)
    _say(f"case: split pack width mode={split_mode} {dtype} k={k}")
    plan, c = gpu_contract(a, b, [1], [1], mode=split_mode)
    exp = O.OraclePlan(a, b, [1], [1]).execute()
    assert rel_frobenius(c.data, exp) <= TOL


@pytest.mark.parametrize("m", [1, 3])
@pytest.mark.parametrize("k", [16, 80, 300, 1000])
@pytest.mark.parametrize("a_transposed", [False, True])
@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_split_pack_few_rows(forced_split, split_mode, dtype, k, m, a_transposed):
    """A blocks of one to three rows: the single-read pack's workgroup then
    holds fewer rows than it has room for (up to 128 rows at the narrowest
    k), in both memory orders of A (k contiguous, or rows contiguous)."""
    ix_m = Index(((0,),), (m,), IN)
    ix_k = Index(((0,),), (k,), OUT)
    ix_n = Index(((0,),), (40,), OUT)
    legs = [ix_k, ix_m] if a_transposed else [ix_m, ix_k]
    a = BlockSparseTensor.allowed_blocks(legs, (0,), dtype).fill_uniform(71)
    b = BlockSparseTensor.allowed_blocks([ix_k.inverse(), ix_n], (0,), dtype).fill_uniform(72)
    a_ax = [0] if a_transposed else [1]
    _say(f"case: split few rows mode={split_mode} {dtype} m={m} k={k} a_transposed={a_transposed}")
    plan, c = gpu_contract(a, b, a_ax, [0], mode=split_mode)
    exp = O.OraclePlan(a, b, a_ax, [0]).execute()
    assert rel_frobenius(c.data, exp) <= TOL


@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_split_modes_alternate_on_one_plan(forced_split, dtype):
    """One cached plan executed in BF16, int8, BF16 and native in turn: each
    switch lays the split workspace out anew (slice planes of 5 BF16 or 6
    int8 digits, their zero padding) and every result stays within 1e-12 of
    the oracle; repeating a mode gives the same bits."""
    rng = np.random.default_rng(4242)
    a, b, a_ax, b_ax = random_case(rng, dtype=dtype, nq=1, max_rank=3, max_sectors=3, max_deg=40, min_deg=20,
                                   allow_scalar=False)
    plan = ContractionPlan(a, b, a_ax, b_ax)
    raw = int(plan.info.c_raw_size)
    da = capi.DeviceBuffer.from_numpy(a.data)
    db = capi.DeviceBuffer.from_numpy(b.data)
    exp = O.OraclePlan(a, b, a_ax, b_ax).execute()
    got = {}
    for step, mode in enumerate([capi.QLT_MODE_SPLIT_BF16, capi.QLT_MODE_SPLIT_I8, capi.QLT_MODE_SPLIT_BF16,
                                 capi.QLT_MODE_NATIVE, capi.QLT_MODE_SPLIT_I8]):
        _say(f"case: split mode switch {dtype} step {step} mode {mode}")
        dc = capi.DeviceBuffer.from_numpy(np.full(max(1, raw), np.nan, dtype=a.np_dtype))
        plan.execute(a.dtype, mode, da.ptr, db.ptr, dc.ptr)
        capi.check(capi.lib().qlt_device_sync())
        c = plan.output_tensor(a.dtype)
        c.data = dc.to_numpy(a.np_dtype, raw)
        assert per_block_rel_err(c, exp) <= TOL, (step, mode)
        if mode in got:
            np.testing.assert_array_equal(c.data, got[mode])
        got[mode] = c.data.copy()
        dc.close()


@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_split_k_beyond_int8_flush_interval(forced_split, split_mode, dtype):
    """k = 8500 crosses the int8 form's 8192-k flush of its int32 digit
    levels; entries in [0.5, 1) of one sign make every level sum grow as fast
    as it can, so a missed flush would overflow int32."""
    ix_m = Index(((0,),), (70,), IN)
    ix_k = Index(((0,),), (8500,), OUT)
    ix_n = Index(((0,),), (40,), OUT)
    a = BlockSparseTensor.allowed_blocks([ix_m, ix_k], (0,), dtype).fill_uniform(51)
    b = BlockSparseTensor.allowed_blocks([ix_k.inverse(), ix_n], (0,), dtype).fill_uniform(52)
    if dtype == "c128":
        a.data = 0.5 + 0.5 * np.abs(a.data.real) + 1j * (0.5 + 0.5 * np.abs(a.data.imag))
        b.data = 0.5 + 0.5 * np.abs(b.data.real) + 1j * (0.5 + 0.5 * np.abs(b.data.imag))
    else:
        a.data = 0.5 + 0.5 * np.abs(a.data)
        b.data = 0.5 + 0.5 * np.abs(b.data)
    plan, c = gpu_contract(a, b, [1], [0], mode=split_mode)
    exp = O.OraclePlan(a, b, [1], [0]).execute()
    assert rel_frobenius(c.data, exp) <= TOL


@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_split_wide_dynamic_range(forced_split, split_mode, dtype):
    """Rows of A and columns of B scaled by 2^-30 .. 2^30: the per-row and
    per-column power-of-two scaling keeps every output block accurate."""
    ix_m = Index(((0,),), (40,), IN)
    ix_k = Index(((0,),), (90,), OUT)
    ix_n = Index(((0,),), (50,), OUT)
    a = BlockSparseTensor.allowed_blocks([ix_m, ix_k], (0,), dtype).fill_uniform(31)
    b = BlockSparseTensor.allowed_blocks([ix_k.inverse(), ix_n], (0,), dtype).fill_uniform(32)
    rs = 2.0 ** np.linspace(-30, 30, 40)
    cs = 2.0 ** np.linspace(30, -30, 50)
    a.data = (a.data.reshape(40, 90) * rs[:, None]).ravel()
    b.data = (b.data.reshape(90, 50) * cs[None, :]).ravel()
    plan, c = gpu_contract(a, b, [1], [0], mode=split_mode)
    exp = O.OraclePlan(a, b, [1], [0]).execute()
    # Each output row (and column) is an inner product against one scaled row
    # of A (column of B), so with per-row / per-column scaling every row and
    # column is accurate on its own (truncation near 2^-45 of its own scale);
    # block-wide scaling would leave the rows scaled by 2^-30 with errors
    # near 2^-45 * 2^30 ~ 1e-4.
    got = c.data.reshape(40, 50)
    ref = exp.reshape(40, 50)
    for i in range(40):
        assert rel_frobenius(got[i], ref[i]) <= 1e-11, ("row", i)
    for j in range(50):
        assert rel_frobenius(got[:, j], ref[:, j]) <= 1e-11, ("col", j)
    assert rel_frobenius(c.data, exp) <= TOL


@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_split_tiny_and_subnormal_rows(forced_split, split_mode, dtype):
    """Rows of A whose maximum is 2^-1000 (normal) or subnormal (2^-1050,
    2^-1070), and columns of B scaled to 2^-1030: the power-of-two scale of
    such a row (2^(46 - e) for int8 digits, 2^-e for BF16 slices) overflows
    FP64 on its own, so the pack must apply it to each element."""
    ix_m = Index(((0,),), (40,), IN)
    ix_k = Index(((0,),), (90,), OUT)
    ix_n = Index(((0,),), (50,), OUT)
    a = BlockSparseTensor.allowed_blocks([ix_m, ix_k], (0,), dtype).fill_uniform(33)
    b = BlockSparseTensor.allowed_blocks([ix_k.inverse(), ix_n], (0,), dtype).fill_uniform(34)
    rs = np.ones(40)
    rs[0:8] = 2.0 ** -1000
    rs[8:12] = 2.0 ** -1050
    rs[12] = 2.0 ** -1070
    cs = np.ones(50)
    cs[45:] = 2.0 ** -1030
    a.data = (a.data.reshape(40, 90) * rs[:, None]).ravel()
    b.data = (b.data.reshape(90, 50) * cs[None, :]).ravel()
    plan, c = gpu_contract(a, b, [1], [0], mode=split_mode)
    exp = O.OraclePlan(a, b, [1], [0]).execute()
    got = c.data.reshape(40, 50)
    ref = exp.reshape(40, 50)
    assert np.all(np.isfinite(got))
    # rows whose products stay normal: relative accuracy of the row
    for i in list(range(8)) + list(range(13, 40)):
        assert rel_frobenius(got[i, :45], ref[i, :45]) <= 1e-11, ("row", i)
    # subnormal results: both sides round every product to the 2^-1074 grid
    # (the oracle once per product, 90 products per entry)
    tiny = 2.0 ** -1074
    assert np.max(np.abs(got[8:13] - ref[8:13])) <= 256 * tiny
    assert np.max(np.abs(got[:, 45:] - ref[:, 45:])) <= 1e-11 * np.max(np.abs(ref[:, 45:])) + 256 * tiny
    assert rel_frobenius(c.data, exp) <= TOL


def test_split_zero_blocks_and_rows(forced_split, split_mode):
    ix = Index(((0,),), (70,), OUT)
    a = BlockSparseTensor.allowed_blocks([ix.inverse(), ix], (0,)).fill_uniform(41)
    b = BlockSparseTensor.allowed_blocks([ix.inverse(), ix], (0,)).fill_uniform(42)
    a.data = a.data.copy()
    a.data.reshape(70, 70)[5:17, :] = 0.0      # zero rows take exponent 0
    plan, c = gpu_contract(a, b, [1], [0], mode=split_mode)
    exp = O.OraclePlan(a, b, [1], [0]).execute()
    assert rel_frobenius(c.data, exp) <= TOL
    assert np.all(c.data.reshape(70, 70)[5:17, :] == 0.0)


def test_split_beta_and_alpha(forced_split, split_mode):
    rng = np.random.default_rng(5)
    a, b, a_ax, b_ax = random_case(rng, dtype="c128", allow_scalar=False, max_deg=8)
    plan = ContractionPlan(a, b, a_ax, b_ax)
    n = plan.info.c_raw_size
    init = (np.arange(n) % 7 - 3.0) + 1j * (np.arange(n) % 5 - 2.0)
    da, db = capi.DeviceBuffer.from_numpy(a.data), capi.DeviceBuffer.from_numpy(b.data)
    dc = capi.DeviceBuffer.from_numpy(init.astype(np.complex128))
    alpha, beta = 0.5 - 1.5j, 2.0 + 0.25j
    plan.execute("c128", split_mode, da.ptr, db.ptr, dc.ptr, alpha=alpha, beta=beta)
    got = dc.to_numpy(np.complex128)
    exp = alpha * O.OraclePlan(a, b, a_ax, b_ax).execute() + beta * init
    assert rel_frobenius(got, exp) <= TOL


@pytest.mark.parametrize("alpha,beta", [(0.5 - 1.5j, 0.0), (0.0, 2.0 + 0.25j), (1.0, 0.5j)],
                         ids=["complex-alpha-beta0", "alpha0-complex-beta", "real-alpha-imag-beta"])
def test_split_complex_scalar_special_cases(forced_split, split_mode, alpha, beta):
    """The complex split kernels take real scalars; a complex alpha or beta is
    applied on their tiles around them.  beta = 0 must not read C (it holds
    NaN here), alpha = 0 only scales C."""
    rng = np.random.default_rng(8)
    a, b, a_ax, b_ax = random_case(rng, dtype="c128", allow_scalar=False, max_deg=40, min_deg=8)
    plan = ContractionPlan(a, b, a_ax, b_ax)
    n = plan.info.c_raw_size
    if beta == 0:
        init = np.full(n, np.nan + 1j * np.nan)
    else:
        init = (np.arange(n) % 7 - 3.0) + 1j * (np.arange(n) % 5 - 2.0)
    da, db = capi.DeviceBuffer.from_numpy(a.data), capi.DeviceBuffer.from_numpy(b.data)
    dc = capi.DeviceBuffer.from_numpy(init.astype(np.complex128))
    plan.execute("c128", split_mode, da.ptr, db.ptr, dc.ptr, alpha=alpha, beta=beta)
    got = dc.to_numpy(np.complex128)
    exp = alpha * O.OraclePlan(a, b, a_ax, b_ax).execute()
    if beta != 0:
        exp = exp + beta * init
    assert np.all(np.isfinite(got))
    assert rel_frobenius(got, exp) <= TOL


def test_split_real_alpha_beta_complex(forced_split, split_mode):
    """Real alpha and beta on complex128: the int8 kernel's waves then store
    their own output component (beta read back per component)."""
    rng = np.random.default_rng(6)
    a, b, a_ax, b_ax = random_case(rng, dtype="c128", allow_scalar=False, max_deg=40, min_deg=8)
    plan = ContractionPlan(a, b, a_ax, b_ax)
    n = plan.info.c_raw_size
    init = (np.arange(n) % 7 - 3.0) + 1j * (np.arange(n) % 5 - 2.0)
    da, db = capi.DeviceBuffer.from_numpy(a.data), capi.DeviceBuffer.from_numpy(b.data)
    dc = capi.DeviceBuffer.from_numpy(init.astype(np.complex128))
    plan.execute("c128", split_mode, da.ptr, db.ptr, dc.ptr, alpha=0.5, beta=2.0)
    got = dc.to_numpy(np.complex128)
    exp = 0.5 * O.OraclePlan(a, b, a_ax, b_ax).execute() + 2.0 * init
    assert rel_frobenius(got, exp) <= TOL


@pytest.mark.parametrize("u1u1,dtype", [(False, "f64"), (True, "c128")])
def test_split_small_matvec_chain(forced_split, split_mode, u1u1, dtype):
    mv = W.small_matvec(D=64 if u1u1 else 96, dtype=dtype, u1u1=u1u1)
    _say(f"case: split small matvec {mv.name}")
    x = x_ref = mv.psi
    for step, (ops, axes) in enumerate(zip((mv.L, mv.W, mv.W, mv.R), W.MATVEC_AXES)):
        first, second = (ops, x) if step == 0 else (x, ops)
        first_ref, second_ref = (ops, x_ref) if step == 0 else (x_ref, ops)
        plan, c = gpu_contract(first, second, list(axes[0]), list(axes[1]), mode=split_mode)
        ref = O.OraclePlan(first_ref, second_ref, list(axes[0]), list(axes[1]))
        c_ref = ref.output_tensor(dtype)
        c_ref.data = ref.execute()
        err = per_block_rel_err(c, c_ref.data)
        assert err <= TOL, (step, err)
        x, x_ref = c, c_ref


def test_split_default_dispatch_keeps_low_k_native(qlt, split_mode):
    """With the default threshold a k=2 contraction runs natively in split
    mode (bit-identical to the native result)."""
    ix = Index(((0,), (1,)), (30, 40), IN)
    p = Index(((0,),), (2,), OUT)
    a = BlockSparseTensor.allowed_blocks([ix, p], (0,)).fill_uniform(1)
    b = BlockSparseTensor.allowed_blocks([p.inverse(), ix.inverse()], (0,)).fill_uniform(2)
    _, c_native = gpu_contract(a, b, [1], [0])
    _, c_split = gpu_contract(a, b, [1], [0], mode=split_mode)
    np.testing.assert_array_equal(c_native.data, c_split.data)
