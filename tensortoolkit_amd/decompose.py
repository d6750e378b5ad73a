"""Block SVD and QR of a block-sparse tensor on the device.

Mirrors the reference's tensor decompositions:

* ``svd`` -- ``SVD(&t, ldims, lqndiv, trunc_err, Dmin, Dmax, &U, &S, &Vt,
  &err, &D)`` (tensor_manipulation/ten_decomp/ten_svd.h:84-263): the blocks
  are grouped into dense block matrices by (left flux, right flux)
  (GenIdxTenDecompDataBlkMats, ten_decomp_basic.h:68-106; InsertDataBlkMat,
  data_blk_mat.h:39-131), every block matrix is decomposed, the singular
  values are truncated globally (CalcTruncedSVInfo_, ten_svd.h:281-372) and
  U, S, Vt are assembled with a new middle index (CreateSVDResTens_ /
  FillSVDResTens_, ten_svd.h:455-581).
* ``qr`` -- ``QR(&t, ldims, lqndiv, &Q, &R)`` (ten_qr.h:56-293).

Block-matrix assembly, the batched decompositions and the scatter of the
factors into the result blocks run on the device (qlt_copy_blocks_2d,
qlt_block_svd, qlt_block_qr); only metadata and the singular values (for
the truncation) touch the host.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import capi
from .blocksparse import BlockSparseTensor, Index, IN, OUT


class EmptySVDResultError(RuntimeError):
    """No singular value survives (EmptySVDResultError, ten_svd.h:196-203)."""


def _qn_add(a, b):
    return tuple(int(x) + int(y) for x, y in zip(a, b))


def _qn_sub(a, b):
    return tuple(int(x) - int(y) for x, y in zip(a, b))


def tensor_div(t: BlockSparseTensor):
    """Div(t): the divergence of the first block (qltensor_impl.h Div / CalcDiv, index.h:268-285)."""
    nq = len(t.indexes[0].qns[0])
    if t.n_blks == 0:
        return tuple([0] * nq)
    q = tuple([0] * nq)
    for ax, ix in enumerate(t.indexes):
        q = _qn_add(q, tuple(ix.dir * v for v in ix.qns[int(t.coords[0, ax])]))
    return q


@dataclass
class BlockMatrix:
    """TenDecompDataBlkMat (data_blk_mat.h:39-110)."""
    lflux: tuple
    rflux: tuple
    mid_qn: tuple
    rows: int = 0
    cols: int = 0
    row_scts: list = field(default_factory=list)   # (lpart coords, offset, dim)
    col_scts: list = field(default_factory=list)   # (rpart coords, offset, dim)
    elems: list = field(default_factory=list)      # (i, j, block position in t)
    _row_pos: dict = field(default_factory=dict, repr=False)   # lpart coords -> row sector
    _col_pos: dict = field(default_factory=dict, repr=False)

    def insert(self, m, n, blk_pos, lpart, rpart):
        i = self._row_pos.get(lpart)
        if i is None:
            i = self._row_pos[lpart] = len(self.row_scts)
            self.row_scts.append((lpart, self.rows, m))
            self.rows += m
        j = self._col_pos.get(rpart)
        if j is None:
            j = self._col_pos[rpart] = len(self.col_scts)
            self.col_scts.append((rpart, self.cols, n))
            self.cols += n
        self.elems.append((i, j, blk_pos))


def block_matrices(t: BlockSparseTensor, ldims: int, lqndiv) -> list:
    """GenIdxTenDecompDataBlkMats (ten_decomp_basic.h:68-106): blocks in
    ascending block index; a block's left flux is sum over the first `ldims`
    legs of (+qn for OUT, -qn for IN) (CalcDataBlkMatLeftQNFlux, :40-66);
    block matrices keyed by (left flux, right flux) in first-seen order."""
    div = tensor_div(t)
    nb = t.n_blks
    if nb == 0:
        return []
    coords = np.asarray(t.coords, dtype=np.int64).reshape(nb, -1)
    dims = np.asarray(t.dims, dtype=np.int64).reshape(nb, -1)
    lfl = np.zeros((nb, len(div)), dtype=np.int64)
    for ax in range(ldims):
        ix = t.indexes[ax]
        lfl += ix.dir * np.asarray(ix.qns, dtype=np.int64).reshape(ix.nsec, -1)[coords[:, ax]]
    rfl = np.asarray(div, dtype=np.int64)[None, :] - lfl
    ms = np.prod(dims[:, :ldims], axis=1).tolist()
    ns = np.prod(dims[:, ldims:], axis=1).tolist()
    lfl, rfl, co = lfl.tolist(), rfl.tolist(), coords.tolist()
    mats: list[BlockMatrix] = []
    by_flux: dict = {}
    for b in range(nb):
        key = (tuple(lfl[b]), tuple(rfl[b]))
        bm = by_flux.get(key)
        if bm is None:
            bm = by_flux[key] = BlockMatrix(key[0], key[1], _qn_sub(tuple(lqndiv), key[0]))
            mats.append(bm)
        bm.insert(ms[b], ns[b], b, tuple(co[b][:ldims]), tuple(co[b][ldims:]))
    return mats


def _esz(dtype):
    return 16 if dtype == "c128" else 8


def _dt(dtype):
    return capi.QLT_C128 if dtype == "c128" else capi.QLT_F64


def _copy_2d(dtype, rects, src_ptr, dst_ptr, stream=None):
    """rects: list of (rows, cols, src_off, src_ld, dst_off, dst_ld)."""
    if not rects:
        return
    a = np.asarray(rects, dtype=np.int64).reshape(-1, 6)
    cols = [np.ascontiguousarray(a[:, k]) for k in range(6)]
    capi.check(capi.lib().qlt_copy_blocks_2d(_dt(dtype), len(a), capi.i64(cols[0]), capi.i64(cols[1]), src_ptr,
                                             capi.i64(cols[2]), capi.i64(cols[3]), dst_ptr, capi.i64(cols[4]),
                                             capi.i64(cols[5]), stream))


def _assemble(t: BlockSparseTensor, mats, t_ptr, stream):
    """Dense block matrices (row-major, concatenated) on the device
    (RawDataGenDenseDataBlkMat_, raw_data_operations.h:550-575)."""
    offs, total = [], 0
    for bm in mats:
        offs.append(total)
        total += bm.rows * bm.cols
    buf = capi.DeviceBuffer(max(1, total) * _esz(t.dtype))
    capi.check(capi.lib().qlt_memset_zero(buf.ptr, max(1, total) * _esz(t.dtype), stream))
    rects = []
    for bm, mo in zip(mats, offs):
        for i, j, b in bm.elems:
            _, ro, m = bm.row_scts[i]
            _, co, n = bm.col_scts[j]
            rects.append((m, n, int(t.offsets[b]), n, mo + ro * bm.cols + co, bm.cols))
    _copy_2d(t.dtype, rects, t_ptr, buf.ptr, stream)
    return buf, offs


def _tensor_from_blocks(indexes, blocks, dtype):
    """blocks: list of (coords, dims, payload) -> tensor with blocks in
    ascending block index (DataBlkInsert order) plus the payloads in that order."""
    t = BlockSparseTensor(indexes, np.zeros((0, len(indexes)), np.int64), np.zeros((0, len(indexes)), np.int64),
                          dtype=dtype)
    if not blocks:
        return t, []
    coords = np.array([b[0] for b in blocks], dtype=np.int64)
    strides = np.ones(len(indexes), dtype=np.int64)
    for j in range(len(indexes) - 2, -1, -1):
        strides[j] = strides[j + 1] * indexes[j + 1].nsec
    order = np.argsort((coords * strides).sum(axis=1), kind="stable")
    blocks = [blocks[k] for k in order]
    t = BlockSparseTensor(indexes, [b[0] for b in blocks], [b[1] for b in blocks], dtype=dtype)
    return t, [b[2] for b in blocks]


def truncate(svals, trunc_err, dmin, dmax):
    """CalcTruncedSVInfo_ (ten_svd.h:281-372).  svals: list of per-matrix
    descending singular values.  Returns (kept count per matrix, actual
    truncation error, D).  Zero singular values are never counted."""
    nm = len(svals)
    lens = [len(sv) for sv in svals]
    s = np.concatenate([np.asarray(sv, dtype=np.float64) for sv in svals]) if sum(lens) else np.zeros(0)
    mi = np.repeat(np.arange(nm), lens)
    nz = s != 0.0
    s, mi = s[nz], mi[nz]
    total_n = len(s)
    if total_n <= dmin:
        return np.bincount(mi, minlength=nm).tolist(), 0.0, total_n
    # descending, ties in (matrix, position) order; c[k] = sum of the k
    # largest squares, summed in that order
    order = np.argsort(-s, kind="stable")
    sq = s[order] * s[order]
    c = np.concatenate([[0.0], np.cumsum(sq)])
    total = c[-1]
    target = total * (1.0 - trunc_err)
    # the smallest kd >= dmin whose kept weight exceeds the target, at most
    # min(total_n, dmax)
    limit = min(total_n, dmax)
    over = np.nonzero(c[dmin:limit + 1] > target)[0]
    kd = dmin + int(over[0]) if len(over) else limit
    err = 1.0 - c[kd] / total
    return np.bincount(mi[order[:kd]], minlength=nm).tolist(), (float(err) if err > 0 else 0.0), kd


@dataclass
class SVDResult:
    u: BlockSparseTensor
    s: BlockSparseTensor
    vt: BlockSparseTensor
    trunc_err: float
    D: int
    sweeps: int
    svals: list          # per block matrix, all singular values (descending)


def svd(t: BlockSparseTensor, ldims: int, lqndiv, trunc_err: float, dmin: int, dmax: int,
        t_dev: "capi.DeviceBuffer | None" = None, stream=None, to_host: bool = True) -> SVDResult:
    """Truncated block SVD (SVD, ten_svd.h:241-263).  The U, S, Vt data are
    kept on the device (``.dev``) and, with ``to_host``, also copied to the
    host (``.data``; otherwise U.data and Vt.data are None)."""
    assert t.rank >= 2 and 1 <= ldims < t.rank and dmin <= dmax and trunc_err >= 0
    mats = block_matrices(t, ldims, lqndiv)
    if t_dev is None:
        t_dev = capi.DeviceBuffer.from_numpy(t.data)
    dtype, esz = t.dtype, _esz(t.dtype)
    a_buf, a_offs = _assemble(t, mats, t_dev.ptr, stream)
    rows = np.array([bm.rows for bm in mats], dtype=np.int64)
    cols = np.array([bm.cols for bm in mats], dtype=np.int64)
    r = np.minimum(rows, cols)
    u_offs = np.concatenate([[0], np.cumsum(rows * r)[:-1]]).astype(np.int64)
    s_offs = np.concatenate([[0], np.cumsum(r)[:-1]]).astype(np.int64)
    vt_offs = np.concatenate([[0], np.cumsum(r * cols)[:-1]]).astype(np.int64)
    u_buf = capi.DeviceBuffer(max(1, int((rows * r).sum())) * esz)
    s_buf = capi.DeviceBuffer(max(1, int(r.sum())) * 8)
    vt_buf = capi.DeviceBuffer(max(1, int((r * cols).sum())) * esz)
    sweeps = C.c_int32()
    a_offs_arr = np.asarray(a_offs, dtype=np.int64)
    capi.check(capi.lib().qlt_block_svd(_dt(dtype), len(mats), capi.i64(rows), capi.i64(cols), a_buf.ptr,
                                        capi.i64(a_offs_arr), u_buf.ptr, capi.i64(u_offs), s_buf.ptr,
                                        capi.i64(s_offs), vt_buf.ptr, capi.i64(vt_offs), C.byref(sweeps), stream))
    s_all = s_buf.to_numpy(np.float64, int(r.sum()))
    svals = [s_all[s_offs[i]:s_offs[i] + r[i]] for i in range(len(mats))]
    kept, err, D = truncate(svals, trunc_err, dmin, dmax)
    if D == 0:
        raise EmptySVDResultError("SVD failed: empty result (no block compatible with lqndiv)")
    # middle index: one sector per block matrix with kept values, in matrix order
    live = [i for i in range(len(mats)) if kept[i] > 0]
    mid_out = Index(tuple(mats[i].mid_qn for i in live), tuple(kept[i] for i in live), OUT)
    mid_in = mid_out.inverse()
    u_blocks, vt_blocks, s_blocks = [], [], []
    for coor, i in enumerate(live):
        bm, k = mats[i], kept[i]
        for lpart, ro, m in bm.row_scts:
            u_blocks.append((list(lpart) + [coor], None, (m, k, int(u_offs[i]) + ro * int(r[i]), int(r[i]))))
        for rpart, co, n in bm.col_scts:
            vt_blocks.append(([coor] + list(rpart), None, (k, n, int(vt_offs[i]) + co, int(cols[i]))))
        s_blocks.append(([coor, coor], [k, k], svals[i][:k]))
    # block shapes follow the indexes' sector degeneracies
    u_idx = list(t.indexes[:ldims]) + [mid_out]
    vt_idx = [mid_in] + list(t.indexes[ldims:])
    s_idx = [mid_in, mid_out]

    def with_dims(idx, blocks):
        return [(c, [idx[a].degs[c[a]] for a in range(len(idx))], p) for c, _, p in blocks]

    U, u_pay = _tensor_from_blocks(u_idx, with_dims(u_idx, u_blocks), dtype)
    Vt, vt_pay = _tensor_from_blocks(vt_idx, with_dims(vt_idx, vt_blocks), dtype)
    S, s_pay = _tensor_from_blocks(s_idx, s_blocks, "f64")
    U.dev = capi.DeviceBuffer(max(1, U.raw_size) * esz)
    Vt.dev = capi.DeviceBuffer(max(1, Vt.raw_size) * esz)
    _copy_2d(dtype, [(m, k, src, ld, int(U.offsets[b]), k) for b, (m, k, src, ld) in enumerate(u_pay)],
             u_buf.ptr, U.dev.ptr, stream)
    _copy_2d(dtype, [(k, n, src, ld, int(Vt.offsets[b]), n) for b, (k, n, src, ld) in enumerate(vt_pay)],
             vt_buf.ptr, Vt.dev.ptr, stream)
    sdata = np.zeros(S.raw_size, dtype=np.float64)
    for b, sv in enumerate(s_pay):
        k = len(sv)
        sdata[S.offsets[b]:S.offsets[b] + k * k].reshape(k, k)[np.arange(k), np.arange(k)] = sv
    S.data = sdata
    S.dev = capi.DeviceBuffer.from_numpy(sdata if sdata.size else np.zeros(1))
    if to_host:
        U.data = U.dev.to_numpy(U.np_dtype, U.raw_size)
        Vt.data = Vt.dev.to_numpy(Vt.np_dtype, Vt.raw_size)
    else:
        capi.check(capi.lib().qlt_stream_sync(stream))
    return SVDResult(U, S, Vt, err, D, int(sweeps.value), svals)


def qr(t: BlockSparseTensor, ldims: int, lqndiv, t_dev: "capi.DeviceBuffer | None" = None, stream=None):
    """Block QR (QR, ten_qr.h:137-159): Q carries t's first `ldims` legs and a
    new OUT middle index with one sector of dim min(rows, cols) per block
    matrix; R carries the inverse middle index and the remaining legs."""
    assert t.rank >= 2 and 1 <= ldims < t.rank
    mats = block_matrices(t, ldims, lqndiv)
    if t_dev is None:
        t_dev = capi.DeviceBuffer.from_numpy(t.data)
    dtype, esz = t.dtype, _esz(t.dtype)
    a_buf, a_offs = _assemble(t, mats, t_dev.ptr, stream)
    rows = np.array([bm.rows for bm in mats], dtype=np.int64)
    cols = np.array([bm.cols for bm in mats], dtype=np.int64)
    k = np.minimum(rows, cols)
    q_offs = np.concatenate([[0], np.cumsum(rows * k)[:-1]]).astype(np.int64)
    r_offs = np.concatenate([[0], np.cumsum(k * cols)[:-1]]).astype(np.int64)
    q_buf = capi.DeviceBuffer(max(1, int((rows * k).sum())) * esz)
    r_buf = capi.DeviceBuffer(max(1, int((k * cols).sum())) * esz)
    a_offs_arr = np.asarray(a_offs, dtype=np.int64)
    capi.check(capi.lib().qlt_block_qr(_dt(dtype), len(mats), capi.i64(rows), capi.i64(cols), a_buf.ptr,
                                       capi.i64(a_offs_arr), q_buf.ptr, capi.i64(q_offs), r_buf.ptr,
                                       capi.i64(r_offs), stream))
    mid_out = Index(tuple(bm.mid_qn for bm in mats), tuple(int(x) for x in k), OUT)
    mid_in = mid_out.inverse()
    q_idx = list(t.indexes[:ldims]) + [mid_out]
    r_idx = [mid_in] + list(t.indexes[ldims:])
    q_blocks, r_blocks = [], []
    for coor, bm in enumerate(mats):
        kk = int(k[coor])
        for lpart, ro, m in bm.row_scts:
            c = list(lpart) + [coor]
            q_blocks.append((c, [q_idx[a].degs[c[a]] for a in range(len(c))],
                             (m, kk, int(q_offs[coor]) + ro * kk, kk)))
        for rpart, co, n in bm.col_scts:
            c = [coor] + list(rpart)
            r_blocks.append((c, [r_idx[a].degs[c[a]] for a in range(len(c))],
                             (kk, n, int(r_offs[coor]) + co, int(cols[coor]))))
    Q, q_pay = _tensor_from_blocks(q_idx, q_blocks, dtype)
    R, r_pay = _tensor_from_blocks(r_idx, r_blocks, dtype)
    Q.dev = capi.DeviceBuffer(max(1, Q.raw_size) * esz)
    R.dev = capi.DeviceBuffer(max(1, R.raw_size) * esz)
    _copy_2d(dtype, [(m, kk, src, ld, int(Q.offsets[b]), kk) for b, (m, kk, src, ld) in enumerate(q_pay)],
             q_buf.ptr, Q.dev.ptr, stream)
    _copy_2d(dtype, [(kk, n, src, ld, int(R.offsets[b]), n) for b, (kk, n, src, ld) in enumerate(r_pay)],
             r_buf.ptr, R.dev.ptr, stream)
    Q.data = Q.dev.to_numpy(Q.np_dtype, Q.raw_size)
    R.data = R.dev.to_numpy(R.np_dtype, R.raw_size)
    return Q, R


def block_svd_dense(mats, dtype="f64"):
    """Batched SVD of a list of dense host matrices through qlt_block_svd
    (test and measurement helper).  Returns [(U, s, Vt)], sweeps."""
    npd = np.complex128 if dtype == "c128" else np.float64
    rows = np.array([a.shape[0] for a in mats], dtype=np.int64)
    cols = np.array([a.shape[1] for a in mats], dtype=np.int64)
    r = np.minimum(rows, cols)
    a_offs = np.concatenate([[0], np.cumsum(rows * cols)[:-1]]).astype(np.int64)
    u_offs = np.concatenate([[0], np.cumsum(rows * r)[:-1]]).astype(np.int64)
    s_offs = np.concatenate([[0], np.cumsum(r)[:-1]]).astype(np.int64)
    vt_offs = np.concatenate([[0], np.cumsum(r * cols)[:-1]]).astype(np.int64)
    a = capi.DeviceBuffer.from_numpy(np.concatenate([np.asarray(m, dtype=npd).ravel() for m in mats]))
    esz = _esz(dtype)
    u = capi.DeviceBuffer(max(1, int((rows * r).sum())) * esz)
    s = capi.DeviceBuffer(max(1, int(r.sum())) * 8)
    vt = capi.DeviceBuffer(max(1, int((r * cols).sum())) * esz)
    sweeps = C.c_int32()
    capi.check(capi.lib().qlt_block_svd(_dt(dtype), len(mats), capi.i64(rows), capi.i64(cols), a.ptr,
                                        capi.i64(a_offs), u.ptr, capi.i64(u_offs), s.ptr, capi.i64(s_offs), vt.ptr,
                                        capi.i64(vt_offs), C.byref(sweeps), None))
    U = u.to_numpy(npd, int((rows * r).sum()))
    S = s.to_numpy(np.float64, int(r.sum()))
    V = vt.to_numpy(npd, int((r * cols).sum()))
    out = []
    for i in range(len(mats)):
        out.append((U[u_offs[i]:u_offs[i] + rows[i] * r[i]].reshape(rows[i], r[i]),
                    S[s_offs[i]:s_offs[i] + r[i]],
                    V[vt_offs[i]:vt_offs[i] + r[i] * cols[i]].reshape(r[i], cols[i])))
    return out, int(sweeps.value)


def block_qr_dense(mats, dtype="f64"):
    """Batched QR of a list of dense host matrices through qlt_block_qr."""
    npd = np.complex128 if dtype == "c128" else np.float64
    rows = np.array([a.shape[0] for a in mats], dtype=np.int64)
    cols = np.array([a.shape[1] for a in mats], dtype=np.int64)
    k = np.minimum(rows, cols)
    a_offs = np.concatenate([[0], np.cumsum(rows * cols)[:-1]]).astype(np.int64)
    q_offs = np.concatenate([[0], np.cumsum(rows * k)[:-1]]).astype(np.int64)
    r_offs = np.concatenate([[0], np.cumsum(k * cols)[:-1]]).astype(np.int64)
    a = capi.DeviceBuffer.from_numpy(np.concatenate([np.asarray(m, dtype=npd).ravel() for m in mats]))
    esz = _esz(dtype)
    q = capi.DeviceBuffer(max(1, int((rows * k).sum())) * esz)
    rr = capi.DeviceBuffer(max(1, int((k * cols).sum())) * esz)
    capi.check(capi.lib().qlt_block_qr(_dt(dtype), len(mats), capi.i64(rows), capi.i64(cols), a.ptr,
                                       capi.i64(a_offs), q.ptr, capi.i64(q_offs), rr.ptr, capi.i64(r_offs), None))
    Q = q.to_numpy(npd, int((rows * k).sum()))
    R = rr.to_numpy(npd, int((k * cols).sum()))
    return [(Q[q_offs[i]:q_offs[i] + rows[i] * k[i]].reshape(rows[i], k[i]),
             R[r_offs[i]:r_offs[i] + k[i] * cols[i]].reshape(k[i], cols[i])) for i in range(len(mats))]
