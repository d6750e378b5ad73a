"""Block-sparse tensor metadata and the contraction plan, for Python callers.

Mirrors what the reference keeps on the host for a ``QLTensor<ElemT,QNT>``
(include/qlten/qltensor/qltensor.h:61-660) and its BlockSparseDataTensor
(include/qlten/qltensor/blk_spar_data_ten/blk_spar_data_ten.h:53-595):
indexes (QN sectors + direction), the list of stored blocks in ascending
block index, their degeneracy shapes and raw-data offsets.  The arithmetic
runs in libqlten_amd through the C ABI.
"""
from __future__ import annotations

import ctypes as C
import itertools
from dataclasses import dataclass, field

import numpy as np

from . import capi

IN, OUT, NDIR = -1, 1, 0
GOLDEN = np.uint64(0x9E3779B97F4A7C15)


@dataclass(frozen=True)
class Index:
    """QN sectors (qn tuple, degeneracy) with a direction (index.h:46-229)."""

    qns: tuple          # tuple of QN tuples, e.g. ((-1,), (0,), (1,)) or ((0, 0), (1, 1))
    degs: tuple         # degeneracy of each sector
    dir: int            # IN = -1, OUT = 1

    def __post_init__(self):
        assert len(self.qns) == len(self.degs)
        assert self.dir in (IN, OUT, NDIR)

    @property
    def nsec(self) -> int:
        return len(self.qns)

    @property
    def dim(self) -> int:
        return int(sum(self.degs))

    def inverse(self) -> "Index":
        """InverseIndex (index.h:240-245)."""
        return Index(self.qns, self.degs, -self.dir)

    def sector_offsets(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.degs)[:-1]]).astype(np.int64)


def u1_index(charges, degs, direction) -> Index:
    return Index(tuple((int(q),) for q in charges), tuple(int(d) for d in degs), direction)


def splitmix_uniform(n: int, seed: int, first: int = 0) -> np.ndarray:
    """Element i = 2u - 1, u = (SplitMix64(seed + (first+i+1) * golden) >> 11) * 2^-53.

    The same generator as qlt_fill_uniform (device) and ora_fill_uniform (C).
    """
    i = np.arange(first + 1, first + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return -1.0 + 2.0 * ((z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53)


class BlockSparseTensor:
    """Metadata of one block-sparse tensor plus its raw data on the host.

    ``coords`` / ``dims`` are (n_blks, rank) int64 arrays in ascending block
    index (row-major over ``blk_shape``); ``offsets`` are raw-data offsets
    (prefix sums of block sizes, as DataBlksInsert computes them,
    data_blk_operations.h:182-208).  ``data`` is a float64 or complex128
    numpy array of ``raw_size`` elements, or None.
    """

    def __init__(self, indexes, coords, dims, dtype="f64", data=None, offsets=None):
        self.indexes = list(indexes)
        self.rank = len(self.indexes)
        self.dtype = dtype
        self.blk_shape = np.array([ix.nsec for ix in self.indexes], dtype=np.int64)
        shape = (-1, self.rank) if self.rank > 0 else (0, 0)   # a rank-0 tensor keeps no blocks
        self.coords = np.ascontiguousarray(np.asarray(coords, dtype=np.int64).reshape(shape))
        self.dims = np.ascontiguousarray(np.asarray(dims, dtype=np.int64).reshape(shape))
        sizes = np.prod(self.dims, axis=1) if len(self.dims) else np.zeros(0, dtype=np.int64)
        if offsets is None:
            offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]) if len(sizes) else np.zeros(0)
        self.offsets = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
        self.raw_size = int(sizes.sum()) if len(sizes) else 0
        self.data = data

    # ---- construction ---------------------------------------------------
    @classmethod
    def allowed_blocks(cls, indexes, div, dtype="f64"):
        """Blocks with CalcDiv(indexes, coords) == div, in row-major coordinate
        order (QLTensor::Random, qltensor_impl.h:374-403; CalcDiv, index.h:268-285)."""
        indexes = list(indexes)
        rank = len(indexes)
        div = np.asarray(div, dtype=np.int64).reshape(-1)
        nq = len(indexes[0].qns[0])
        qarr = [np.asarray(ix.qns, dtype=np.int64).reshape(ix.nsec, nq) * ix.dir for ix in indexes]
        # enumerate the first rank-1 axes, then solve for the last one
        if rank == 1:
            sel = [c for c in range(indexes[0].nsec) if np.array_equal(qarr[0][c], div)]
            coords = np.array(sel, dtype=np.int64).reshape(-1, 1)
        else:
            grids = np.meshgrid(*[np.arange(ix.nsec) for ix in indexes[:-1]], indexing="ij")
            head = np.stack([g.reshape(-1) for g in grids], axis=1)          # row-major order
            partial = np.zeros((head.shape[0], nq), dtype=np.int64)
            for ax in range(rank - 1):
                partial += qarr[ax][head[:, ax]]
            need = div[None, :] - partial                                     # contribution of last axis
            last = qarr[-1]
            rows, cols = [], []
            # match every head row with every last-axis sector giving `need`
            lookup = {}
            for c in range(last.shape[0]):
                lookup.setdefault(tuple(last[c]), []).append(c)
            for r in range(head.shape[0]):
                for c in lookup.get(tuple(need[r]), ()):
                    rows.append(r)
                    cols.append(c)
            rows = np.asarray(rows, dtype=np.int64)
            cols = np.asarray(cols, dtype=np.int64)
            coords = np.concatenate([head[rows], cols[:, None]], axis=1) if len(rows) else \
                np.zeros((0, rank), dtype=np.int64)
        degs = [np.asarray(ix.degs, dtype=np.int64) for ix in indexes]
        dims = np.stack([degs[ax][coords[:, ax]] for ax in range(rank)], axis=1) if len(coords) else \
            np.zeros((0, rank), dtype=np.int64)
        return cls(indexes, coords, dims, dtype=dtype)

    def fill_uniform(self, seed: int) -> "BlockSparseTensor":
        """Uniform(-1,1) entries in raw order (complex: re, im interleaved)."""
        n = self.raw_size * (2 if self.dtype == "c128" else 1)
        x = splitmix_uniform(n, seed)
        self.data = x.view(np.complex128) if self.dtype == "c128" else x
        return self

    # ---- queries --------------------------------------------------------
    @property
    def n_blks(self) -> int:
        return int(self.coords.shape[0])

    @property
    def blk_idx(self) -> np.ndarray:
        strides = np.ones(self.rank, dtype=np.int64)
        for j in range(self.rank - 2, -1, -1):
            strides[j] = strides[j + 1] * self.blk_shape[j + 1]
        return (self.coords * strides[None, :]).sum(axis=1) if self.n_blks else np.zeros(0, np.int64)

    @property
    def np_dtype(self):
        return np.complex128 if self.dtype == "c128" else np.float64

    def block(self, i: int) -> np.ndarray:
        size = int(np.prod(self.dims[i]))
        return self.data[self.offsets[i]:self.offsets[i] + size].reshape(tuple(self.dims[i]))

    def desc(self) -> capi.BsdtDesc:
        d = capi.BsdtDesc()
        d.rank = self.rank
        d.n_blks = self.n_blks
        d.blk_shape = capi.i64(self.blk_shape)
        d.blk_coords = capi.i64(self.coords)
        d.blk_dims = capi.i64(self.dims)
        d.blk_offsets = capi.i64(self.offsets)
        return d

    def to_dense(self) -> np.ndarray:
        """Dense array over the full index dimensions (test helper)."""
        full = np.zeros(tuple(ix.dim for ix in self.indexes), dtype=self.np_dtype)
        soffs = [ix.sector_offsets() for ix in self.indexes]
        for i in range(self.n_blks):
            sl = tuple(slice(soffs[ax][self.coords[i, ax]], soffs[ax][self.coords[i, ax]] + self.dims[i, ax])
                       for ax in range(self.rank))
            full[sl] = self.block(i)
        return full

    def nbytes(self) -> int:
        return self.raw_size * (16 if self.dtype == "c128" else 8)


def saved_axes(rank: int, ctrct) -> list:
    """TenCtrctGenSavedAxesSet (data_blk_operations.h:268-296): ascending non-contracted axes."""
    return [i for i in range(rank) if i not in ctrct]


def contiguous_axes(rank: int, start: int, size: int):
    """Contracted and saved axes of the contiguous path
    (contract_contiguous_axes.h:342-353): contracted (start+i)%rank, saved (end+i)%rank."""
    ctr = [(start + i) % rank for i in range(size)]
    end = (start + size) % rank
    sav = [(end + i) % rank for i in range(rank - size)]
    return ctr, sav


class ContractionPlan:
    """A plan built on the device by qlt_plan_build."""

    def __init__(self, a: BlockSparseTensor, b: BlockSparseTensor, a_ctrct, b_ctrct,
                 a_saved=None, b_saved=None, stream=None):
        capi.init()
        self.a, self.b = a, b
        self.a_ctrct = list(a_ctrct)
        self.b_ctrct = list(b_ctrct)
        self.a_saved = saved_axes(a.rank, a_ctrct) if a_saved is None else list(a_saved)
        self.b_saved = saved_axes(b.rank, b_ctrct) if b_saved is None else list(b_saved)
        self._ax = [np.ascontiguousarray(np.asarray(v, dtype=np.int32)) for v in
                    (self.a_ctrct, self.b_ctrct, self.a_saved, self.b_saved)]
        da, db = a.desc(), b.desc()
        handle = C.c_void_p()
        capi.check(capi.lib().qlt_plan_build(
            C.byref(da), C.byref(db), len(self.a_ctrct), capi.i32(self._ax[0]), capi.i32(self._ax[1]),
            len(self.a_saved), capi.i32(self._ax[2]), len(self.b_saved), capi.i32(self._ax[3]),
            stream, C.byref(handle)))
        self.handle = handle
        info = capi.PlanInfo()
        capi.check(capi.lib().qlt_plan_get_info(self.handle, C.byref(info)))
        self.info = info
        self.c_indexes = [a.indexes[i] for i in self.a_saved] + [b.indexes[i] for i in self.b_saved]

    def tasks(self) -> np.ndarray:
        out = np.zeros(self.info.n_tasks, dtype=capi.TASK_DTYPE)
        if len(out):
            capi.check(capi.lib().qlt_plan_get_tasks(self.handle, out.ctypes.data_as(C.c_void_p)))
        return out

    def c_blocks(self):
        n, r = self.info.n_c_blks, self.info.c_rank
        idx = np.zeros(n, np.int64)
        coords = np.zeros((n, r), np.int64)
        dims = np.zeros((n, r), np.int64)
        offs = np.zeros(n, np.int64)
        if n:
            capi.check(capi.lib().qlt_plan_get_c_blocks(self.handle, capi.i64(idx), capi.i64(coords),
                                                        capi.i64(dims), capi.i64(offs)))
        return idx, coords, dims, offs

    def output_tensor(self, dtype: str) -> BlockSparseTensor:
        _, coords, dims, offs = self.c_blocks()
        return BlockSparseTensor(self.c_indexes, coords, dims, dtype=dtype, offsets=offs)

    def prepare(self, dtype: str, mode: int) -> None:
        capi.check(capi.lib().qlt_plan_prepare(self.handle, _dt(dtype), mode))

    def execute(self, dtype: str, mode: int, a_ptr, b_ptr, c_ptr, alpha=1.0, beta=0.0, stream=None):
        al = np.array([np.real(alpha), np.imag(alpha)], dtype=np.float64)
        be = np.array([np.real(beta), np.imag(beta)], dtype=np.float64)
        capi.check(capi.lib().qlt_contract_execute(self.handle, _dt(dtype), mode, a_ptr, b_ptr, c_ptr,
                                                   capi.f64(al), capi.f64(be), stream))

    def flops(self, dtype: str) -> float:
        """Sum over tasks of c*m*k*n, c = 2 real / 8 complex (tensor_op_cost.h:384-406)."""
        return (8.0 if dtype == "c128" else 2.0) * float(self.info.sum_mkn)

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value is not None:
            capi.lib().qlt_plan_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _dt(dtype: str) -> int:
    return capi.QLT_C128 if dtype == "c128" else capi.QLT_F64


def contract(a: BlockSparseTensor, b: BlockSparseTensor, axes, mode: int = capi.QLT_MODE_NATIVE,
             a_saved=None, b_saved=None) -> BlockSparseTensor:
    """Contract(A, B, {a_axes, b_axes}) on the device; host arrays in and out."""
    assert a.dtype == b.dtype
    plan = ContractionPlan(a, b, axes[0], axes[1], a_saved, b_saved)
    c = plan.output_tensor(a.dtype)
    if plan.info.c_is_scalar:
        c.raw_size = int(plan.info.c_raw_size)
    da = capi.DeviceBuffer.from_numpy(a.data)
    db = capi.DeviceBuffer.from_numpy(b.data)
    esz = 16 if a.dtype == "c128" else 8
    dc = capi.DeviceBuffer(max(1, c.raw_size) * esz)
    plan.execute(a.dtype, mode, da.ptr, db.ptr, dc.ptr)
    capi.check(capi.lib().qlt_device_sync())
    c.data = dc.to_numpy(c.np_dtype, c.raw_size)
    plan.close()
    for buf in (da, db, dc):
        buf.close()
    return c
