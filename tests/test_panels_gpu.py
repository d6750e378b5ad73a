"""Row panels (qlt_set_row_panels): the narrow output blocks (n <= 4) that
share their coordinates on A's saved axes run as one unit that reads each A
element once for all of them.

Checks, through the C ABI:
* the W contractions of the matvec chains build panels (plan info), and
  their output with panels equals the output of the per-block row kernel bit
  for bit (alpha = 1, beta = 0), in all four element types, and lies within
  1e-12 (FP64 types) or float rounding (single types) of the oracle;
* complex alpha / beta against the oracle's alpha A B + beta C0;
* the split modes run the panels natively (the same bits);
* an output block reading more than four A' columns, or an A with more than
  four saved axes, stays on the row kernel (no panels) and is still right;
* accumulating into a shifted output layout (qlt_plan_set_c_offsets): the
  panels read the output offsets at run time.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tensortoolkit_amd import capi, workloads as W
from tensortoolkit_amd.blocksparse import IN, OUT, BlockSparseTensor, ContractionPlan, Index
from tests.helpers import per_block_rel_err

pytestmark = pytest.mark.gpu

CODES = {"f64": (capi.QLT_F64, np.float64), "c128": (capi.QLT_C128, np.complex128),
         "f32": (capi.QLT_F32, np.float32), "c64": (capi.QLT_C64, np.complex64)}
WIDE = {"f32": "f64", "c64": "c128", "f64": "f64", "c128": "c128"}


def _say(msg):
    print(msg, flush=True)


def build(first, second, a_ax, b_ax, panels):
    lib = capi.lib()
    old = lib.qlt_get_row_panels()
    capi.check(lib.qlt_set_row_panels(1 if panels else 0))
    try:
        return ContractionPlan(first, second, a_ax, b_ax)
    finally:
        capi.check(lib.qlt_set_row_panels(old))


def execute(plan, a, b, dtype, mode=capi.QLT_MODE_NATIVE, alpha=1.0, beta=0.0, c_init=None, c_size=None):
    code, npt = CODES[dtype]
    raw = int(plan.info.c_raw_size) if c_size is None else c_size
    da = capi.DeviceBuffer.from_numpy(a.data.astype(npt))
    db = capi.DeviceBuffer.from_numpy(b.data.astype(npt))
    init = np.full(max(1, raw), np.nan, dtype=npt) if c_init is None else c_init.astype(npt)
    dc = capi.DeviceBuffer.from_numpy(init)
    al = np.array([np.real(alpha), np.imag(alpha)])
    be = np.array([np.real(beta), np.imag(beta)])
    capi.check(capi.lib().qlt_contract_execute(plan.handle, code, mode, da.ptr, db.ptr, dc.ptr, capi.f64(al),
                                               capi.f64(be), None))
    capi.check(capi.lib().qlt_device_sync())
    out = dc.to_numpy(npt, raw)
    for buf in (da, db, dc):
        buf.close()
    return out


_ops = {}


def w_operands(name, step):
    """The operands of W contraction `step` (1 or 2) of a matvec chain: the
    previous output's structure filled uniform(-1, 1), and the MPO tensor."""
    key = (name, step)
    if key not in _ops:
        mv = {"u1_D2048": lambda: W.config1(), "u1u1_D4096": lambda: W.config3(),
              "u1u1_small": lambda: W.small_matvec(D=160, dtype="c128", u1u1=True)}[name]()
        x = mv.psi
        for s in range(step):
            op = (mv.L, mv.W)[s]
            first, second = (op, x) if s == 0 else (x, op)
            ax = W.MATVEC_AXES[s]
            x = O.OraclePlan(first, second, list(ax[0]), list(ax[1])).output_tensor(mv.dtype)
        x.fill_uniform(9000 + step)
        ax = W.MATVEC_AXES[step]
        _ops[key] = (x, mv.W, list(ax[0]), list(ax[1]))
    return _ops[key]


def rounded(t, dtype):
    """A copy of tensor t holding the values of `dtype` (float-rounded for the
    single types), stored in t's own FP64 type for the oracle."""
    c = BlockSparseTensor(t.indexes, t.coords, t.dims, t.dtype, offsets=t.offsets)
    c.data = t.data.astype(CODES[dtype][1]).astype(t.np_dtype)
    return c


CASES = [("u1_D2048", 1, "f64"), ("u1_D2048", 2, "f32"), ("u1u1_small", 1, "c128"), ("u1u1_small", 2, "c64"),
         ("u1u1_D4096", 1, "c128"), ("u1u1_D4096", 2, "c128")]


@pytest.mark.parametrize("name,step,dtype", CASES)
def test_panels_match_row_kernel_and_oracle(qlt, name, step, dtype):
    first, second, a_ax, b_ax = w_operands(name, step)
    if WIDE[dtype] != first.dtype:
        pytest.skip("element type of another chain")
    _say(f"case: panels {name} W step {step} {dtype}")
    a, b = rounded(first, dtype), rounded(second, dtype)
    on = build(a, b, a_ax, b_ax, True)
    off = build(a, b, a_ax, b_ax, False)
    assert on.info.n_panel_tiles > 0
    assert off.info.n_panel_tiles == 0
    got_on = execute(on, a, b, dtype)
    got_off = execute(off, a, b, dtype)
    np.testing.assert_array_equal(got_on, got_off)
    exp = O.OraclePlan(a, b, a_ax, b_ax).execute()
    c = on.output_tensor(a.dtype)
    c.data = got_on.astype(c.np_dtype)
    tol = 1e-12 if dtype in ("f64", "c128") else 1.2e-7
    assert per_block_rel_err(c, exp) <= tol
    # the split modes run the row work natively
    if dtype in ("f64", "c128"):
        for mode in (capi.QLT_MODE_SPLIT_BF16, capi.QLT_MODE_SPLIT_I8):
            np.testing.assert_array_equal(execute(on, a, b, dtype, mode), got_on)
    on.close()
    off.close()


@pytest.mark.parametrize("dtype", ["c128", "f64", "c64", "f32"])
def test_panels_alpha_beta(qlt, dtype):
    # single-precision storage: float-rounded inputs and C, FP64 arithmetic,
    # one rounding per output element (relative error below 2^-23 per element)
    wide = WIDE[dtype]
    single = dtype in ("c64", "f32")
    first, second, a_ax, b_ax = w_operands("u1u1_small", 1)
    a, b = rounded(first, "c64" if single else "c128"), rounded(second, "c64" if single else "c128")
    if wide == "f64":
        a.data, b.data = a.data.real.copy(), b.data.real.copy()
        a.dtype = b.dtype = "f64"
    _say(f"case: panels alpha/beta {dtype}")
    plan = build(a, b, a_ax, b_ax, True)
    assert plan.info.n_panel_tiles > 0
    raw = int(plan.info.c_raw_size)
    rng = np.random.default_rng(77)
    c0 = rng.uniform(-1, 1, raw) + (1j * rng.uniform(-1, 1, raw) if wide == "c128" else 0)
    c0 = c0.astype(CODES[dtype][1]).astype(CODES[wide][1])
    alpha, beta = ((0.7 - 0.3j), (0.2 + 0.5j)) if wide == "c128" else (0.7, -1.3)
    got = execute(plan, a, b, dtype, alpha=alpha, beta=beta, c_init=c0).astype(CODES[wide][1])
    exp = alpha * O.OraclePlan(a, b, a_ax, b_ax).execute() + beta * c0
    assert np.linalg.norm(got - exp) <= (2e-7 if single else 1e-12) * np.linalg.norm(exp)
    off = build(a, b, a_ax, b_ax, False)
    ref = execute(off, a, b, dtype, alpha=alpha, beta=beta, c_init=c0).astype(CODES[wide][1])
    assert np.max(np.abs(got - ref)) <= (1e-6 if single else 1e-14) * np.max(np.abs(ref))
    plan.close()
    off.close()


def test_panels_shifted_output_layout(qlt):
    """qlt_plan_set_c_offsets moves every output block by `pad` elements; the
    panels write there and leave the pad untouched."""
    a, b, a_ax, b_ax = w_operands("u1u1_small", 2)
    plan = build(a, b, a_ax, b_ax, True)
    assert plan.info.n_panel_tiles > 0
    raw = int(plan.info.c_raw_size)
    n_c = int(plan.info.n_c_blks)
    offs = np.zeros(n_c, dtype=np.int64)
    capi.check(capi.lib().qlt_plan_get_c_blocks(plan.handle, None, None, None, capi.i64(offs)))
    pad = 37
    capi.check(capi.lib().qlt_plan_set_c_offsets(plan.handle, capi.i64(offs + pad)))
    got = execute(plan, a, b, "c128", c_size=raw + pad)
    assert np.all(np.isnan(got[:pad]))
    exp = O.OraclePlan(a, b, a_ax, b_ax).execute()
    assert np.linalg.norm(got[pad:] - exp) <= 1e-12 * np.linalg.norm(exp)
    plan.close()


def test_no_panel_when_more_than_four_columns(qlt):
    """Output blocks whose tasks read 8 A' columns stay on the row kernel."""
    ix_r = Index(((0,), (1,)), (1500, 700), IN)
    ix_k = Index(((0,),), (8,), OUT)
    ix_n = Index(((0,), (1,)), (2, 3), OUT)
    a = BlockSparseTensor.allowed_blocks([ix_r, ix_k, ix_n.inverse()], (0,), "c128").fill_uniform(5)
    b = BlockSparseTensor.allowed_blocks([ix_k.inverse(), ix_n], (0,), "c128").fill_uniform(6)
    _say("case: panels fallback k=8")
    plan = build(a, b, [1], [0], True)
    assert plan.info.n_panel_tiles == 0
    got = execute(plan, a, b, "c128")
    c = plan.output_tensor("c128")
    c.data = got
    assert per_block_rel_err(c, O.OraclePlan(a, b, [1], [0]).execute()) <= 1e-12
    plan.close()


def test_no_panel_beyond_four_saved_axes(qlt):
    """An A with five saved axes stays on the row kernel."""
    ix = Index(((0,), (1,)), (3, 2), IN)
    ix_k = Index(((0,), (1,)), (1, 1), OUT)
    legs = [ix, ix.inverse(), ix, ix.inverse(), ix, ix_k]
    a = BlockSparseTensor.allowed_blocks(legs, (0,), "f64").fill_uniform(7)
    b = BlockSparseTensor.allowed_blocks([ix_k.inverse(), Index(((0,),), (2,), OUT)], (0,), "f64").fill_uniform(8)
    _say("case: panels fallback five saved axes")
    plan = build(a, b, [5], [0], True)
    assert plan.info.n_panel_tiles == 0
    got = execute(plan, a, b, "f64")
    c = plan.output_tensor("f64")
    c.data = got
    assert per_block_rel_err(c, O.OraclePlan(a, b, [5], [0]).execute()) <= 1e-12
    plan.close()


def test_panel_switch_rejects_other_values(qlt):
    lib = capi.lib()
    assert lib.qlt_set_row_panels(2) != capi.QLT_OK
    assert lib.qlt_get_row_panels() in (0, 1)
