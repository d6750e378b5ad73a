"""CPU pin of the reference Lanczos (oracle/lanczos.py): on a small Hermitian
matvec its lowest Ritz value after enough steps equals the lowest eigenvalue
of the dense H built column by column, and with few steps it is an upper
bound (Rayleigh-Ritz)."""
import numpy as np
import pytest

from oracle import lanczos as OL
from tensortoolkit_amd import workloads as W
from tests.test_workloads import dense_heff


@pytest.mark.parametrize("u1u1,dtype,D", [(False, "f64", 10), (True, "c128", 16)])
def test_lanczos_matches_dense_eigensolver(u1u1, dtype, D):
    mv = W.hermitian_matvec(W.small_matvec(D=D, dtype=dtype, u1u1=u1u1))
    H = dense_heff(mv)
    e0 = np.linalg.eigvalsh(H)[0]
    n = H.shape[0]
    theta, gs, a, b, _ = OL.lanczos(mv, mv.psi.data.copy(), steps=n)
    assert abs(theta - e0) <= 1e-10 * abs(e0)
    r = H @ gs - theta * gs
    assert np.linalg.norm(r) <= 1e-8 * abs(e0)
    theta5, *_ = OL.lanczos(mv, mv.psi.data.copy(), steps=5)
    assert theta5 >= e0 - 1e-12
