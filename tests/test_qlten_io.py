"""The .qlten on-disk format on the host (tensortoolkit_amd/qlten_io.py):
round trips, the stored hashes, and malformed input."""
import io

import numpy as np
import pytest

from tensortoolkit_amd import qlten_io as Q, workloads as W
from tensortoolkit_amd.blocksparse import BlockSparseTensor, Index, IN, OUT


def _roundtrip(t):
    buf = io.BytesIO()
    Q.write(t, buf)
    raw = buf.getvalue()
    r = Q.read(io.BytesIO(raw), t.dtype)
    assert r.rank == t.rank
    for x, y in zip(r.indexes, t.indexes):
        assert (x.qns, x.degs, x.dir) == (y.qns, y.degs, y.dir)
    # blocks come back in ascending block index
    order = np.argsort(t.blk_idx, kind="stable")
    np.testing.assert_array_equal(r.coords, t.coords[order])
    for k, i in enumerate(order):
        s = int(np.prod(t.dims[i]))
        np.testing.assert_array_equal(r.data[r.offsets[k]:r.offsets[k] + s], t.data[t.offsets[i]:t.offsets[i] + s])
    # writing what was read reproduces the bytes
    buf2 = io.BytesIO()
    Q.write(r, buf2)
    assert buf2.getvalue() == raw
    return raw


def test_roundtrip_u1_real():
    a, b, _ = W.config0()
    raw = _roundtrip(a)
    # the text header ends with the block coordinates, then one line break and
    # the raw data, then one line break
    assert raw.endswith(a.data.tobytes() + b"\n")


def test_roundtrip_u1u1_complex():
    mv = W.small_matvec(D=64, dtype="c128", u1u1=True)
    _roundtrip(mv.psi)
    _roundtrip(mv.W)


def test_roundtrip_blocks_out_of_order():
    a, _, _ = W.config0()
    perm = np.arange(a.n_blks)[::-1]
    t = BlockSparseTensor(a.indexes, a.coords[perm], a.dims[perm], dtype=a.dtype)
    t.data = np.concatenate([a.data[a.offsets[i]:a.offsets[i] + int(np.prod(a.dims[i]))] for i in perm])
    r = Q.read(io.BytesIO(_roundtrip(t)), t.dtype)
    np.testing.assert_array_equal(r.data, a.data)       # ascending block index, as a


def test_scalar_and_empty():
    s = BlockSparseTensor([], np.zeros((0, 0), np.int64), np.zeros((0, 0), np.int64), dtype="c128")
    s.data = np.array([2.5 - 1j])
    buf = io.BytesIO()
    Q.write(s, buf)
    raw = buf.getvalue()
    assert raw == b"0\n0\n" + np.array([2.5 - 1j]).tobytes() + b"\n"
    r = Q.read(io.BytesIO(raw), "c128")
    assert r.rank == 0 and r.data.tolist() == [2.5 - 1j]
    ix = Index(((0,), (1,)), (2, 3), OUT)
    e = BlockSparseTensor([ix, ix.inverse()], np.zeros((0, 2), np.int64), np.zeros((0, 2), np.int64))
    e.data = np.zeros(0)
    _roundtrip(e)


def test_hash_closed_forms():
    # rotations leave 0 and all-ones fixed; 1 rotated left by 31 is 2^31
    assert Q.qn_hash((0,)) == 0 and Q.qn_hash((-1,)) == 2**64 - 1 and Q.qn_hash((1,)) == 2**31
    # U1U1: (2^30) + (2^30) 2^31 = 2^30 + 2^61, rotated left by 10: 2^40 + 2^7
    assert Q.qn_hash((0, 0)) == 2**40 + 2**7
    assert Q.sector_hash((1,), 3) == 2**31 ^ 3
    # the index hash depends on the direction and on the sector order
    ix = Index(((0,), (1,)), (2, 3), OUT)
    assert Q.index_hash(ix) != Q.index_hash(ix.inverse())
    assert Q.index_hash(ix) != Q.index_hash(Index(((1,), (0,)), (3, 2), OUT))
    # VecHasher of an empty list: P5 + (0 ^ P5)
    assert Q.vec_hash([]) == (2 * 2870177450012600261) % 2**64


def test_malformed_input_is_rejected():
    a, _, _ = W.config0()
    buf = io.BytesIO()
    Q.write(a, buf)
    raw = buf.getvalue()
    with pytest.raises(ValueError):
        Q.read(io.BytesIO(raw[:-100]), a.dtype)                    # truncated data
    lines = raw.split(b"\n")
    lines[3] = b"12345"                                             # corrupt the first QN hash
    with pytest.raises(ValueError):
        Q.read(io.BytesIO(b"\n".join(lines)), a.dtype, qn_len=1)
