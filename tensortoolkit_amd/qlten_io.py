"""The reference's on-disk tensor format (``.qlten``), read and written on the host.

Layout (QLTensor::StreamWrite, qltensor_impl.h:823-830; Index::StreamWrite,
index.h:182-188; QNSector::StreamWrite, qnsct.h:111-115; U1QN / U1U1QN
StreamWrite, u1qn.h:131-133 / u1u1qn.h:137-139; BlockSparseDataTensor::
StreamWrite, blk_spar_data_ten.h:809-826; RawDataWrite_,
raw_data_operations.h:612-633), every number in decimal on its own line:

    rank
    per index:  n_sectors
                per sector:  qn value(s), qn hash, degeneracy, sector hash
                direction (IN = -1, OUT = 1), dimension, index hash
    n_blocks
    per block (ascending block index): its coordinates
    raw data: n elements of float64 or complex128 (host byte order), then '\\n'

The stored hashes are those of the reference's QN and Index types: U1QN is
its value rotated left by 31 bits (u1qn.h:139-152), U1U1QN packs
(v1 + 2^30) + (v2 + 2^30) * 2^31 and rotates left by 10 (u1u1qn.h:151-173), a
sector is qn_hash ^ degeneracy (qnsct.h:121-125), an index is
VecHasher(sectors) ^ std::hash<int>(direction) (index.h:225-228, vec_hash.h),
all in 64-bit wrapping arithmetic.  Readers keep the hashes they read, as the
reference does.  Used to load real tensors as fixtures and by the tests; the
C++ host layer reads and writes the same bytes (``operator<<`` / ``operator>>``
in ``include/qlten/qltensor/qltensor.h``).
"""
from __future__ import annotations

import io

import numpy as np

from .blocksparse import BlockSparseTensor, Index

_M = (1 << 64) - 1
_P1, _P2, _P5 = 11400714785074694791, 14029467366897019727, 2870177450012600261


def _rotl(x: int, r: int) -> int:
    x &= _M
    return ((x << r) | (x >> (64 - r))) & _M


def qn_hash(qn: tuple) -> int:
    if len(qn) == 1:
        return _rotl(qn[0] & _M, 31)
    if len(qn) == 2:
        seg = 1 << 30
        h = (((qn[0] & _M) + seg) + ((qn[1] & _M) + seg) * (2 * seg)) & _M
        return _rotl(h, 10)
    raise ValueError("only U(1) and U(1)xU(1) quantum numbers are supported")


def sector_hash(qn: tuple, deg: int) -> int:
    return qn_hash(qn) ^ (deg & _M)


def vec_hash(hashes) -> int:
    h = _P5
    for x in hashes:
        h = (h + x * _P2) & _M
        h = _rotl(h, 31)
        h = (h * _P1) & _M
    return (h + (len(hashes) ^ _P5)) & _M


def index_hash(ix: Index) -> int:
    return vec_hash([sector_hash(q, d) for q, d in zip(ix.qns, ix.degs)]) ^ (ix.dir & _M)


def write(t: BlockSparseTensor, f) -> None:
    """Write ``t`` to the binary file object ``f``."""
    out = io.StringIO()
    out.write(f"{t.rank}\n")
    for ix in t.indexes:
        out.write(f"{ix.nsec}\n")
        for q, d in zip(ix.qns, ix.degs):
            for v in q:
                out.write(f"{int(v)}\n")
            out.write(f"{qn_hash(q)}\n{int(d)}\n{sector_hash(q, int(d))}\n")
        out.write(f"{int(ix.dir)}\n{ix.dim}\n{index_hash(ix)}\n")
    if t.rank == 0:
        # a scalar keeps no block, only its one raw element
        out.write("0\n")
        f.write(out.getvalue().encode())
        data = np.ascontiguousarray(t.data, dtype=t.np_dtype).reshape(-1)[:1]
        if data.size == 0:
            data = np.zeros(1, dtype=t.np_dtype)
        f.write(data.tobytes())
        f.write(b"\n")
        return
    order = np.argsort(np.asarray(t.blk_idx), kind="stable") if t.n_blks else np.zeros(0, np.int64)
    out.write(f"{t.n_blks}\n")
    for i in order:
        for c in t.coords[i]:
            out.write(f"{int(c)}\n")
    f.write(out.getvalue().encode())
    # raw data in ascending block index, each block row-major
    for i in order:
        size = int(np.prod(t.dims[i]))
        blk = np.ascontiguousarray(t.data[t.offsets[i]:t.offsets[i] + size], dtype=t.np_dtype)
        f.write(blk.tobytes())
    f.write(b"\n")


class _Tokens:
    def __init__(self, buf: bytes):
        self.buf, self.pos = buf, 0

    def next_int(self) -> int:
        b, n = self.buf, len(self.buf)
        while self.pos < n and b[self.pos] in b" \t\r\n":
            self.pos += 1
        start = self.pos
        while self.pos < n and b[self.pos] not in b" \t\r\n":
            self.pos += 1
        if start == self.pos:
            raise ValueError("truncated .qlten header")
        return int(b[start:self.pos])


def read(f, dtype: str, qn_len: int | None = None) -> BlockSparseTensor:
    """Read a tensor of element type ``dtype`` ("f64" or "c128") from the
    binary file object ``f``.  ``qn_len`` is 1 for U1QN and 2 for U1U1QN (the
    reader's template parameter in the reference); None detects it from the
    first sector's stored QN hash.  Blocks
    come back in ascending block index."""
    buf = f.read()
    tk = _Tokens(buf)
    rank = tk.next_int()
    indexes = []
    for _ in range(rank):
        n = tk.next_int()
        qns, degs = [], []
        for _ in range(n):
            if qn_len is None:
                save_pos = tk.pos
                v0, v1, v2 = tk.next_int(), tk.next_int(), tk.next_int()
                # U1U1 first: a U1 sector (v, hash(v), deg, ...) with v = 0 has
                # v1 = 0 = hash(0), which a U1U1 (0, 0, ...) sector also has
                if (v2 & _M) == qn_hash((v0, v1)):
                    qn_len = 2
                elif (v1 & _M) == qn_hash((v0,)):
                    qn_len = 1
                else:
                    raise ValueError("the stored quantum-number hash matches neither U1QN nor U1U1QN")
                tk.pos = save_pos
            qn = tuple(tk.next_int() for _ in range(qn_len))
            stored = tk.next_int()
            if (stored & _M) != qn_hash(qn):
                raise ValueError(f"quantum number {qn}: stored hash {stored} differs from the computed one")
            degs.append(tk.next_int())
            tk.next_int()                                  # stored sector hash
            qns.append(qn)
        direction = tk.next_int()
        tk.next_int()                                      # dimension
        tk.next_int()                                      # stored index hash
        indexes.append(Index(tuple(qns), tuple(degs), direction))
    n_blks = tk.next_int()
    coords = np.array([[tk.next_int() for _ in range(rank)] for _ in range(n_blks)],
                      dtype=np.int64).reshape(n_blks, rank)
    dims = np.array([[indexes[a].degs[coords[i, a]] for a in range(rank)] for i in range(n_blks)],
                    dtype=np.int64).reshape(n_blks, rank)
    t = BlockSparseTensor(indexes, coords, dims, dtype=dtype)
    n = 1 if rank == 0 else t.raw_size
    start = tk.pos + 1                                     # the line break before the data
    esz = 16 if dtype == "c128" else 8
    raw = buf[start:start + n * esz]
    if len(raw) != n * esz:
        raise ValueError("truncated .qlten raw data")
    t.data = np.frombuffer(raw, dtype=t.np_dtype).copy()
    return t


def save(path: str, t: BlockSparseTensor) -> None:
    with open(path, "wb") as f:
        write(t, f)


def load(path: str, dtype: str, qn_len: int | None = None) -> BlockSparseTensor:
    with open(path, "rb") as f:
        return read(f, dtype, qn_len)
