"""CPU pins of the block-decomposition oracle (oracle/decomp.py)."""
import numpy as np
import pytest

from oracle import decomp as OD
from tensortoolkit_amd.blocksparse import BlockSparseTensor, Index, IN, OUT


def u1(charges, degs, d):
    return Index(tuple((q,) for q in charges), tuple(degs), d)


def rand_tensor(seed, dtype="f64", rank=4, max_deg=6):
    rng = np.random.default_rng(seed)
    idx = []
    for ax in range(rank):
        ch = list(range(-2, 3))
        degs = [int(x) for x in rng.integers(1, max_deg + 1, size=len(ch))]
        idx.append(u1(ch, degs, IN if ax < rank // 2 else OUT))
    return BlockSparseTensor.allowed_blocks(idx, (0,), dtype=dtype).fill_uniform(seed + 1)


def test_known_answer_2x2():
    # tests/package_consumer/smoke_gpu.cu:100-117
    ix0, ix1 = u1([0], [2], IN), u1([0], [2], OUT)
    t = BlockSparseTensor([ix0, ix1], [[0, 0]], [[2, 2]], data=np.array([1.0, 2.0, 3.0, 4.0]))
    out, err, D, svals = OD.svd(t, 1, 0.0, 1, 10)
    np.testing.assert_allclose(svals[0], [5.464985704219043, 0.365966190626257], rtol=1e-14)
    assert D == 2 and err == 0.0


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_block_singular_values_are_those_of_the_matricisation(seed, dtype):
    t = rand_tensor(seed, dtype)
    _, _, _, svals = OD.svd(t, 2, 0.0, 1, 10**6)
    got = np.sort(np.concatenate(svals))[::-1]
    d = t.to_dense()
    M = d.reshape(int(np.prod(d.shape[:2])), -1)
    ref = np.linalg.svd(M, compute_uv=False)
    ref = ref[: len(got)]
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * ref[0])
    # every singular value of the matricisation beyond the block ones is zero
    assert np.all(np.linalg.svd(M, compute_uv=False)[len(got):] < 1e-12 * ref[0])


def test_truncation_closed_form():
    # weights 9, 4, 1 (total 14): with trunc_err 0.1 the target is 12.6, so
    # D grows from Dmin = 1 (9 <= 12.6) to 2 (13 > 12.6); error = 1 - 13/14.
    kept, err, D = OD.truncation([np.array([3.0, 1.0]), np.array([2.0])], 0.1, 1, 10)
    assert D == 2 and kept == [1, 1]
    assert abs(err - 1.0 / 14.0) < 1e-15
    # Dmax caps; Dmin floors; zeros are never kept
    assert OD.truncation([np.array([3.0, 2.0, 1.0])], 0.0, 1, 2)[2] == 2
    assert OD.truncation([np.array([3.0, 2.0, 1.0])], 0.9, 3, 5)[2] == 3
    assert OD.truncation([np.array([3.0, 0.0])], 0.0, 1, 5)[2] == 1


@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_untruncated_reconstruction_is_exact(dtype):
    t = rand_tensor(7, dtype)
    M, err, D = OD.truncated_reconstruction(t, 2, 0.0, 1, 10**6)
    d = t.to_dense()
    np.testing.assert_allclose(M, d.reshape(M.shape), atol=1e-12)
    assert 0.0 <= err < 1e-14   # 1 - kept/total with every value kept


def test_truncation_weight_matches_discarded_norm():
    t = rand_tensor(11, "f64", max_deg=8)
    M, err, D = OD.truncated_reconstruction(t, 2, 1e-2, 1, 10**6)
    d = t.to_dense().reshape(M.shape)
    # squared-norm convention: err = |T - T_D|^2 / |T|^2 (ten_svd.h:223-224)
    assert abs(np.linalg.norm(d - M) ** 2 / np.linalg.norm(d) ** 2 - err) < 1e-12


@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_qr_blocks_reconstruct(dtype):
    """Q R equals the assembled block matrix of its left flux, Q has
    orthonormal columns and R is upper triangular; the block matrices hold
    every block of the tensor once (their squared norms add up to the
    tensor's), so a QR of the wrong matrix, or a dropped block, fails."""
    t = rand_tensor(3, dtype)
    groups = OD.group(t, 2)
    res = OD.qr_blocks(t, 2)
    assert [lf for lf, _, _ in res] == [g[0] for g in groups]
    total = 0.0
    for (lf, Q, R), (_, rows, cols, blks) in zip(res, groups):
        A, roff, coff = OD.dense_matrix(t, rows, cols, blks)
        assert Q.shape == (A.shape[0], min(A.shape)) and R.shape == (min(A.shape), A.shape[1])
        np.testing.assert_allclose(Q @ R, A, rtol=0, atol=1e-13 * max(1.0, np.abs(A).max()))
        assert np.allclose(Q.conj().T @ Q, np.eye(Q.shape[1]), atol=1e-13)
        assert np.all(np.tril(R, -1) == 0)
        total += np.linalg.norm(A) ** 2
        # each block sits at its row / column sector offsets
        for lp, rp, b in blks:
            m, n = rows[lp], cols[rp]
            np.testing.assert_array_equal(A[roff[lp]:roff[lp] + m, coff[rp]:coff[rp] + n],
                                          t.block(b).reshape(m, n))
    assert abs(total - np.linalg.norm(t.data) ** 2) <= 1e-12 * total
