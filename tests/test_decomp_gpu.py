"""GPU parity of the batched block SVD / QR (qlt_block_svd, qlt_block_qr) and
of the tensor-level decompositions (tensortoolkit_amd.decompose) against
numpy's LAPACK and the decomposition oracle.

Bounds (FP64): singular values within 1e-13 * sigma_max * sqrt(n) of LAPACK's,
|U diag(S) Vt - A| <= 1e-12 |A|, max |U^H U - I| and |Vt Vt^H - I| <= 1e-12;
QR: Q and R equal LAPACK's (same reflector convention) to 1e-12 on full-rank
matrices.
"""
import numpy as np
import pytest

from oracle import decomp as OD
from tensortoolkit_amd import decompose as DEC
from tests.test_decomp_oracle import rand_tensor

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def rand_mat(rng, m, n, cplx, graded=False):
    a = rng.standard_normal((m, n)) + (1j * rng.standard_normal((m, n)) if cplx else 0)
    if graded:
        u, _, vt = np.linalg.svd(a, full_matrices=False)
        a = (u * 10.0 ** (-np.linspace(0, 12, min(m, n)))) @ vt
    return a


def check_svd(a, u, s, vt, recon_tol=1e-12, orth_tol=1e-12):
    k = min(a.shape)
    ref = np.linalg.svd(a, compute_uv=False)
    smax = max(ref[0], 1e-300)
    assert np.all(np.diff(s) <= 0)
    assert np.max(np.abs(s - ref)) <= 1e-13 * smax * np.sqrt(max(a.shape)) + 1e-300
    assert np.linalg.norm((u * s) @ vt - a) <= recon_tol * max(np.linalg.norm(a), 1e-300)
    nz = s > 1e-13 * smax
    uu = u[:, nz].conj().T @ u[:, nz]
    vv = vt[nz].conj() @ vt[nz].T
    assert np.max(np.abs(uu - np.eye(nz.sum()))) <= orth_tol
    assert np.max(np.abs(vv - np.eye(nz.sum()))) <= orth_tol
    assert u.shape == (a.shape[0], k) and vt.shape == (k, a.shape[1])


def test_svd_known_answer(qlt):
    _say("case: svd 2x2 known answer")
    (u, s, vt), = DEC.block_svd_dense([np.array([[1.0, 2.0], [3.0, 4.0]])])[0]
    np.testing.assert_allclose(s, [5.464985704219043, 0.365966190626257], rtol=1e-14)


@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_svd_batched_shapes(qlt, dtype):
    rng = np.random.default_rng(42)
    shapes = [(1, 1), (1, 7), (9, 1), (2, 2), (17, 5), (5, 17), (33, 33), (64, 40), (40, 64), (100, 100),
              (130, 31), (31, 130)]
    mats = [rand_mat(rng, m, n, dtype == "c128") for m, n in shapes]
    _say(f"case: svd batched {dtype} shapes {shapes}")
    out, sweeps = DEC.block_svd_dense(mats, dtype)
    assert sweeps <= 40
    for a, (u, s, vt) in zip(mats, out):
        check_svd(a, u, s, vt)


@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_svd_large_blocks(qlt, dtype):
    """Block matrices of the end-to-end configuration's sizes, up to 1024 x
    1024 (square, tall, wide), in one batched call.  The rotations of the
    one-sided block Jacobi accumulate rounding with the width, and a column
    pair counts as converged at a cosine of max(16, rows) eps: the
    reconstruction and orthogonality bounds are 1e-12 up to 256 columns and
    grow linearly beyond (measured at 1024 x 1024: 2.0e-12 |A| and 2.4e-12;
    the singular values keep their bound)."""
    rng = np.random.default_rng(1024 + (dtype == "c128"))
    shapes = [(512, 512), (1024, 1024), (1024, 600), (600, 1024), (777, 513)]
    mats = [rand_mat(rng, m, n, dtype == "c128") for m, n in shapes]
    _say(f"case: svd large blocks {dtype} shapes {shapes}")
    out, sweeps = DEC.block_svd_dense(mats, dtype)
    _say(f"sweeps {sweeps}")
    assert sweeps <= 40
    for a, (u, s, vt) in zip(mats, out):
        grow = max(1.0, min(a.shape) / 256)
        check_svd(a, u, s, vt, recon_tol=1e-12 * grow, orth_tol=1e-12 * grow)


@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_svd_graded_and_rank_deficient(qlt, dtype):
    rng = np.random.default_rng(7)
    cp = dtype == "c128"
    mats = [rand_mat(rng, 120, 120, cp, graded=True), rand_mat(rng, 200, 90, cp, graded=True),
            rand_mat(rng, 60, 150, cp, graded=True)]
    low = rand_mat(rng, 50, 8, cp) @ rand_mat(rng, 8, 70, cp)    # rank 8
    z = rand_mat(rng, 30, 20, cp)
    z[:, 5] = 0.0                                                  # a zero column
    mats += [low, z, np.zeros((6, 4))]
    _say(f"case: svd graded / rank-deficient {dtype}")
    out, sweeps = DEC.block_svd_dense(mats, dtype)
    for a, (u, s, vt) in zip(mats, out):
        if not np.any(a):
            assert np.all(s == 0.0)
            continue
        check_svd(a, u, s, vt)


@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_qr_matches_lapack(qlt, dtype):
    rng = np.random.default_rng(3)
    shapes = [(1, 1), (3, 1), (1, 4), (8, 8), (40, 17), (17, 40), (128, 96), (96, 128)]
    mats = [rand_mat(rng, m, n, dtype == "c128") for m, n in shapes]
    _say(f"case: qr batched {dtype} shapes {shapes}")
    out = DEC.block_qr_dense(mats, dtype)
    for a, (q, r) in zip(mats, out):
        qn, rn = np.linalg.qr(a)
        assert np.max(np.abs(q - qn)) <= 1e-12 * max(1, np.sqrt(a.shape[0]))
        assert np.max(np.abs(r - rn)) <= 1e-12 * np.abs(a).max() * a.shape[0]
        assert np.allclose(np.tril(r, -1), 0)
        assert np.linalg.norm(q @ r - a) <= 1e-13 * np.linalg.norm(a) * np.sqrt(a.shape[0])


def test_qr_zero_column(qlt):
    a = np.random.default_rng(1).standard_normal((10, 6))
    a[:, 2] = 0.0
    (q, r), = DEC.block_qr_dense([a])
    assert np.linalg.norm(q @ r - a) <= 1e-13 * np.linalg.norm(a)
    assert np.max(np.abs(q.T @ q - np.eye(6))) <= 1e-12


@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("dtype", ["f64", "c128"])
@pytest.mark.parametrize("trunc", [0.0, 1e-3])
def test_tensor_svd_matches_oracle(qlt, seed, dtype, trunc):
    t = rand_tensor(100 + seed, dtype, max_deg=7)
    _say(f"case: tensor svd seed={seed} {dtype} trunc={trunc} blocks={t.n_blks}")
    res = DEC.svd(t, 2, (0,), trunc, 1, 10**6)
    ref, err_ref, D_ref, svals_ref = OD.svd(t, 2, trunc, 1, 10**6)
    assert res.D == D_ref
    assert abs(res.trunc_err - err_ref) <= 1e-12
    for got, exp in zip(res.svals, svals_ref):
        np.testing.assert_allclose(got, exp, rtol=0, atol=1e-13 * max(1.0, exp[0]) * 4)
    # U S Vt equals the oracle's truncated reconstruction
    Mref, _, _ = OD.truncated_reconstruction(t, 2, trunc, 1, 10**6)
    u, s, vt = res.u.to_dense(), res.s.to_dense(), res.vt.to_dense()
    ul = u.reshape(-1, u.shape[-1])
    vr = vt.reshape(vt.shape[0], -1)
    M = ul @ s @ vr
    assert np.linalg.norm(M - Mref) <= 1e-12 * np.linalg.norm(Mref)
    # U is an isometry on the kept space
    assert np.max(np.abs(ul.conj().T @ ul - np.eye(ul.shape[1]))) <= 1e-12
    assert res.s.indexes[0].dim == res.D


@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_tensor_qr_reconstructs(qlt, dtype):
    t = rand_tensor(5, dtype, max_deg=7)
    _say(f"case: tensor qr {dtype}")
    Q, R = DEC.qr(t, 2, (0,))
    q = Q.to_dense()
    r = R.to_dense()
    ql = q.reshape(-1, q.shape[-1])
    rr = r.reshape(r.shape[0], -1)
    d = t.to_dense()
    assert np.linalg.norm(ql @ rr - d.reshape(ql.shape[0], -1)) <= 1e-12 * np.linalg.norm(d)
    assert np.max(np.abs(ql.conj().T @ ql - np.eye(ql.shape[1]))) <= 1e-12
