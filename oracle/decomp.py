"""CPU oracle of the block decompositions -- TEST INFRASTRUCTURE ONLY.

A restatement of the reference's tensor SVD and QR with numpy's LAPACK
(numpy.linalg.svd = gesdd, numpy.linalg.qr = geqrf + orgqr/ungqr) as the
per-block-matrix routine, returning dense factor tensors of the same block
structure.  Only ``tests/`` and the end-to-end check of ``bench.py``
import it.

Reference steps restated:
  * block matrices: ten_decomp_basic.h:68-106 (GenIdxTenDecompDataBlkMats),
    data_blk_mat.h:39-131 (row / column sectors in first-seen order);
  * truncation: ten_svd.h:281-372 (CalcTruncedSVInfo_: sort singular values
    descending, keep at least Dmin, grow while the kept weight does not
    exceed (1 - trunc_err) of the total and the dimension stays <= Dmax;
    error = 1 - kept / total; zero singular values are dropped);
  * result tensors: ten_svd.h:455-581, ten_qr.h:186-289.

Pinning: the known-answer singular values of [[1, 2], [3, 4]]
(tests/package_consumer/smoke_gpu.cu:100-117), the singular values of the
densified tensor's matricisation (the union over block matrices), U S Vt
reconstruction, the truncation weight closed form, and for the QR: Q R equal
to the assembled block matrix, orthonormal Q, triangular R, every block at
its sector offsets, in tests/test_decomp_oracle.py.
"""
from __future__ import annotations

import numpy as np


def _flux(t, coords, axes):
    nq = len(t.indexes[0].qns[0])
    q = np.zeros(nq, dtype=np.int64)
    for ax in axes:
        ix = t.indexes[ax]
        q += ix.dir * np.asarray(ix.qns[int(coords[ax])], dtype=np.int64)
    return tuple(int(x) for x in q)


def group(t, ldims):
    """[(lflux, rows: [(lpart, dim)], cols: [(rpart, dim)], blocks: [(lpart, rpart, blk)])] in first-seen order."""
    out = {}
    order = []
    for b in range(t.n_blks):
        co = t.coords[b]
        lf = _flux(t, co, range(ldims))
        if lf not in out:
            out[lf] = ({}, {}, [])
            order.append(lf)
        rows, cols, blks = out[lf]
        lp, rp = tuple(int(x) for x in co[:ldims]), tuple(int(x) for x in co[ldims:])
        rows.setdefault(lp, int(np.prod(t.dims[b][:ldims])))
        cols.setdefault(rp, int(np.prod(t.dims[b][ldims:])))
        blks.append((lp, rp, b))
    return [(lf,) + out[lf] for lf in order]


def dense_matrix(t, rows, cols, blks):
    roff, coff = {}, {}
    acc = 0
    for k, d in rows.items():
        roff[k] = acc
        acc += d
    nr = acc
    acc = 0
    for k, d in cols.items():
        coff[k] = acc
        acc += d
    A = np.zeros((nr, acc), dtype=t.np_dtype)
    for lp, rp, b in blks:
        m, n = rows[lp], cols[rp]
        A[roff[lp]:roff[lp] + m, coff[rp]:coff[rp] + n] = t.block(b).reshape(m, n)
    return A, roff, coff


def truncation(svals, trunc_err, dmin, dmax):
    """Kept count per block matrix, truncation error, D (ten_svd.h:281-372)."""
    allv = sorted(((s, i) for i, sv in enumerate(svals) for s in sv if s != 0.0), key=lambda x: -x[0])
    kept = [0] * len(svals)
    if len(allv) <= dmin:
        for _, i in allv:
            kept[i] += 1
        return kept, 0.0, len(allv)
    w = np.array([s * s for s, _ in allv])
    total = w.sum()
    d = dmin
    kw = w[:d].sum()
    while kw <= total * (1.0 - trunc_err) and d + 1 <= len(allv) and d + 1 <= dmax:
        kw += w[d]
        d += 1
    for _, i in allv[:d]:
        kept[i] += 1
    return kept, max(0.0, 1.0 - kw / total), d


def svd(t, ldims, trunc_err, dmin, dmax):
    """Returns (per block matrix: (lflux, U_kept, s_kept, Vt_kept, roff, coff,
    rows, cols)), err, D, all singular values per block matrix."""
    groups = group(t, ldims)
    res, svals = [], []
    for lf, rows, cols, blks in groups:
        A, roff, coff = dense_matrix(t, rows, cols, blks)
        U, s, Vt = np.linalg.svd(A, full_matrices=False)
        res.append((lf, U, s, Vt, roff, coff, rows, cols))
        svals.append(s)
    kept, err, D = truncation(svals, trunc_err, dmin, dmax)
    out = [(lf, U[:, :k], s[:k], Vt[:k, :], roff, coff, rows, cols)
           for (lf, U, s, Vt, roff, coff, rows, cols), k in zip(res, kept)]
    return out, err, D, svals


def truncated_reconstruction(t, ldims, trunc_err, dmin, dmax):
    """Dense U S Vt of the kept singular triplets, as the dense matricisation
    (prod of left dims) x (prod of right dims)."""
    out, err, D, _ = svd(t, ldims, trunc_err, dmin, dmax)
    lsoff = [ix.sector_offsets() for ix in t.indexes[:ldims]]
    rsoff = [ix.sector_offsets() for ix in t.indexes[ldims:]]
    ldim = int(np.prod([ix.dim for ix in t.indexes[:ldims]]))
    rdim = int(np.prod([ix.dim for ix in t.indexes[ldims:]]))
    full = np.zeros(tuple(ix.dim for ix in t.indexes), dtype=t.np_dtype)
    for lf, U, s, Vt, roff, coff, rows, cols in out:
        M = (U * s) @ Vt
        for lp, m in rows.items():
            for rp, n in cols.items():
                sub = M[roff[lp]:roff[lp] + m, coff[rp]:coff[rp] + n]
                if not np.any(sub):
                    continue
                co = list(lp) + list(rp)
                shape = [t.indexes[a].degs[co[a]] for a in range(t.rank)]
                sl = tuple(slice(t.indexes[a].sector_offsets()[co[a]],
                                 t.indexes[a].sector_offsets()[co[a]] + shape[a]) for a in range(t.rank))
                full[sl] += sub.reshape(shape)
    return full.reshape(ldim, rdim), err, D


def qr_blocks(t, ldims):
    """Per block matrix (lflux, Q, R) with numpy's LAPACK QR."""
    out = []
    for lf, rows, cols, blks in group(t, ldims):
        A, _, _ = dense_matrix(t, rows, cols, blks)
        Q, R = np.linalg.qr(A)
        out.append((lf, Q, R))
    return out
