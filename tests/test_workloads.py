"""CPU checks of the synthetic workloads: the Hermitian matvec really is
Hermitian (dense H built column by column through the oracle), and the
unsymmetrised one is not, so the check has teeth."""
import numpy as np
import pytest

from oracle import oracle as O
from tensortoolkit_amd import workloads as W


def dense_heff(mv):
    psi = mv.psi
    n = psi.raw_size
    plans, cur = [], psi
    for step, axes in enumerate(W.MATVEC_AXES):
        op = (mv.L, mv.W, mv.W, mv.R)[step]
        first, second = (op, cur) if step == 0 else (cur, op)
        p = O.OraclePlan(first, second, list(axes[0]), list(axes[1]))
        plans.append(p)
        cur = p.output_tensor(mv.dtype)
    H = np.zeros((n, n), dtype=psi.np_dtype)
    for k in range(n):
        e = np.zeros(n, dtype=psi.np_dtype)
        e[k] = 1
        plans[0].b.data = e
        x = plans[0].execute()
        for p in plans[1:]:
            p.a.data = x
            x = p.execute()
        H[:, k] = x
    return H


@pytest.mark.parametrize("u1u1,dtype,D", [(False, "f64", 10), (True, "c128", 16)])
def test_hermitian_matvec_is_hermitian(u1u1, dtype, D):
    raw = W.small_matvec(D=D, dtype=dtype, u1u1=u1u1)
    H0 = dense_heff(raw)
    assert np.abs(H0 - H0.conj().T).max() > 1e-3 * np.abs(H0).max()
    H = dense_heff(W.hermitian_matvec(raw))
    assert np.abs(H - H.conj().T).max() <= 1e-14 * np.abs(H).max()


def test_degeneracy_split_needs_one_per_sector():
    with pytest.raises(ValueError):
        W.largest_remainder([1.0, 1.0, 1.0], 2)
    assert W.largest_remainder([1.0, 2.0, 1.0], 8).sum() == 8


def test_config1_environments_are_hermitian():
    """The D = 2048 matvec's L and R are symmetrised: as dense arrays,
    L[a', w, a] = conj L[a, w~, a'] with w~ the conjugate MPO sector (the
    Heisenberg bond sectors -1 and +1 swap, 0 stays), and likewise R."""
    mv = W.config1()
    for t in (mv.L, mv.R):
        d = t.to_dense()
        ix = t.indexes[1]
        offs = ix.sector_offsets()
        perm = np.empty(ix.dim, dtype=np.int64)
        part = W._partner_sector(ix)
        for k in range(ix.nsec):
            for j in range(ix.degs[k]):
                perm[offs[k] + j] = offs[part[k]] + j
        np.testing.assert_array_equal(d.transpose(2, 1, 0)[:, perm, :], np.conj(d))
        assert not np.array_equal(d, d.transpose(2, 1, 0))   # not trivially symmetric
