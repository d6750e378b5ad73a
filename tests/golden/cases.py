"""Deterministic input cases of the golden plan fixtures."""
from tensortoolkit_amd import workloads as W
from tensortoolkit_amd.blocksparse import IN, OUT, BlockSparseTensor, Index


def rank3_u1_pair():
    a, b, ax = W.config0()
    return a, b, list(ax[0]), list(ax[1]), False


def cost_known_answer():
    """tests/test_tensor_manipulation/test_tensor_op_cost.cc:19-42."""
    qn0 = ((0,),)
    lhs = BlockSparseTensor.allowed_blocks([Index(qn0, (2,), OUT), Index(qn0, (3,), OUT)], (0,)).fill_uniform(1)
    rhs = BlockSparseTensor.allowed_blocks([Index(qn0, (3,), IN), Index(qn0, (4,), OUT)], (0,)).fill_uniform(2)
    return lhs, rhs, [1], [0], True


def u1u1_multiblock():
    """Modelled on test_ten_ctrct.cc:1425-1467: sectors (-1,-1) x2, (0,0) x3,
    (1,1) x2; lhs (out,out,out,out), rhs (in,in,out,out), divergence 0;
    lhs axes 2,3 against rhs axes 0,1."""
    qns = ((-1, -1), (0, 0), (1, 1))
    idx_in = Index(qns, (2, 3, 2), IN)
    idx_out = idx_in.inverse()
    lhs = BlockSparseTensor.allowed_blocks([idx_out] * 4, (0, 0)).fill_uniform(11)
    rhs = BlockSparseTensor.allowed_blocks([idx_in, idx_in, idx_out, idx_out], (0, 0)).fill_uniform(12)
    return lhs, rhs, [2, 3], [0, 1], True


def matvec_steps(u1u1):
    mv = W.small_matvec(D=64 if u1u1 else 96, dtype="f64", u1u1=u1u1)
    from oracle import oracle as O
    out = {}
    x = mv.psi
    for step, axes in enumerate(W.MATVEC_AXES):
        op = (mv.L, mv.W, mv.W, mv.R)[step]
        first, second = (op, x) if step == 0 else (x, op)
        name = f"matvec_{'u1u1' if u1u1 else 'u1'}_step{step}"
        out[name] = (first, second, list(axes[0]), list(axes[1]), False)
        p = O.OraclePlan(first, second, list(axes[0]), list(axes[1]))
        x = p.output_tensor("f64")
    return out


def all_cases():
    c = {
        "rank3_u1_pair": rank3_u1_pair(),
        "cost_known_answer": cost_known_answer(),
        "u1u1_multiblock": u1u1_multiblock(),
    }
    c.update(matvec_steps(False))
    c.update(matvec_steps(True))
    return c
