"""qlt_scale_segments_noalloc, the entry point that allocates nothing (its
segment list is staged in a fixed device array of the library), against
qlt_scale_segments, the allocating entry point it varies: two calls in a row,
the second with fewer segments, so that what the first call left in the
fixed array must not be read."""
import ctypes as C

import numpy as np
import pytest

from tensortoolkit_amd import capi

pytestmark = pytest.mark.gpu
MAX_SEGS = 65536


def _segments(rng, n_elems, n_segs):
    """n_segs disjoint (offset, length) pairs inside [0, n_elems), sorted."""
    cuts = np.sort(rng.choice(np.arange(1, n_elems), size=2 * n_segs, replace=False))
    offs = cuts[0::2].astype(np.int64)
    lens = (cuts[1::2] - cuts[0::2]).astype(np.int64)
    return offs, lens


def _scale(fn, dtype, data, offs, lens, s):
    buf = capi.DeviceBuffer.from_numpy(data)
    sc = np.ascontiguousarray(np.asarray(s, dtype=np.float64))
    capi.check(fn(dtype, len(offs), capi.i64(offs), capi.i64(lens), sc.ctypes.data_as(C.POINTER(C.c_double)),
                  buf.ptr, None))
    capi.check(capi.lib().qlt_device_sync())
    return buf.to_numpy(data.dtype)


@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_noalloc_matches_allocating_entry_point(qlt, dtype):
    lib = capi.lib()
    rng = np.random.default_rng(1234 + (dtype == "c128"))
    code = capi.QLT_C128 if dtype == "c128" else capi.QLT_F64
    s = [0.75, -1.25] if dtype == "c128" else [-1.5]
    n_elems = 400_000
    for n_segs in (5000, 37):   # the second call has fewer segments than the first
        print(f"case: scale_segments_noalloc {dtype} n_segs={n_segs}", flush=True)
        if dtype == "c128":
            data = rng.uniform(-1, 1, n_elems) + 1j * rng.uniform(-1, 1, n_elems)
        else:
            data = rng.uniform(-1, 1, n_elems)
        offs, lens = _segments(rng, n_elems, n_segs)
        ref = _scale(lib.qlt_scale_segments, code, data, offs, lens, s)
        got = _scale(lib.qlt_scale_segments_noalloc, code, data, offs, lens, s)
        np.testing.assert_array_equal(got, ref)   # the same kernel: bit for bit
        # and the scaling itself: scaled inside the segments, untouched outside
        exp = data.copy()
        factor = complex(*s) if dtype == "c128" else s[0]
        for o, n in zip(offs, lens):
            exp[o:o + n] *= factor
        inside = np.zeros(n_elems, bool)
        for o, n in zip(offs, lens):
            inside[o:o + n] = True
        np.testing.assert_array_equal(got[~inside], data[~inside])
        np.testing.assert_allclose(got, exp, rtol=1e-15, atol=0)


def test_noalloc_rejects_more_segments_than_its_array_holds(qlt):
    lib = capi.lib()
    n = MAX_SEGS + 1
    offs = np.arange(n, dtype=np.int64)
    lens = np.ones(n, dtype=np.int64)
    buf = capi.DeviceBuffer(8 * n)
    sc = np.array([2.0])
    st = lib.qlt_scale_segments_noalloc(capi.QLT_F64, n, capi.i64(offs), capi.i64(lens),
                                        sc.ctypes.data_as(C.POINTER(C.c_double)), buf.ptr, None)
    assert st == 1   # QLT_ERR_INVALID_ARG
    assert b"maximum of 65536" in lib.qlt_last_error()
    # exactly the maximum is accepted
    data = np.arange(n, dtype=np.float64)
    got = _scale(lib.qlt_scale_segments_noalloc, capi.QLT_F64, data, offs[:MAX_SEGS], lens[:MAX_SEGS], [2.0])
    np.testing.assert_array_equal(got[:MAX_SEGS], 2.0 * data[:MAX_SEGS])
    assert got[MAX_SEGS] == data[MAX_SEGS]
