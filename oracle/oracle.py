"""CPU oracle -- TEST INFRASTRUCTURE ONLY.

Python front end of ``oracle.c``, the plain-C restatement of the reference's
plan builder (data_blk_operations.h:411-578) and executor
(global_operations.h:895-1080).  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` import this module, as the checker
or as the timed CPU baseline; the product path never does.

Pinning: see the header of oracle.c.  Everything here is also checked against
brute force and dense numpy arithmetic in tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes as C
import glob
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "liboracle.so"
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)


class OraBsdt(C.Structure):
    _fields_ = [("rank", C.c_int32), ("n_blks", C.c_int64), ("blk_shape", _i64p),
                ("blk_coords", _i64p), ("blk_dims", _i64p), ("blk_offsets", _i64p)]


_lib = None
_blas = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run tensortoolkit_amd.build first")
        h = C.CDLL(str(LIB_PATH))
        h.ora_plan_build.argtypes = [C.POINTER(OraBsdt), C.POINTER(OraBsdt), C.c_int32, _i32p, _i32p,
                                     C.c_int32, _i32p, C.c_int32, _i32p]
        h.ora_plan_build.restype = C.c_void_p
        h.ora_plan_free.argtypes = [C.c_void_p]
        for name in ("ora_plan_n_tasks", "ora_plan_n_c", "ora_plan_c_raw_size"):
            getattr(h, name).argtypes = [C.c_void_p]
            getattr(h, name).restype = C.c_int64
        h.ora_plan_get_tasks.argtypes = [C.c_void_p, _i64p, C.POINTER(C.c_double)]
        h.ora_plan_get_c_blocks.argtypes = [C.c_void_p, _i64p, _i64p, _i64p, _i64p]
        h.ora_execute.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        h.ora_execute.restype = C.c_int
        h.ora_set_blas.argtypes = [C.c_void_p, C.c_void_p]
        h.ora_set_threads.argtypes = [C.c_int]
        h.ora_part_hash.argtypes = [_i64p, _i32p, C.c_int32]
        h.ora_part_hash.restype = C.c_uint64
        h.ora_fill_uniform.argtypes = [C.POINTER(C.c_double), C.c_int64, C.c_uint64, C.c_int64]
        _lib = h
    return _lib


def openblas_path() -> str | None:
    """scipy's bundled OpenBLAS (exports scipy_cblas_* names)."""
    try:
        import scipy  # noqa: F401
    except ImportError:
        return None
    import scipy as sp
    base = Path(sp.__file__).resolve().parent.parent / "scipy.libs"
    hits = sorted(glob.glob(str(base / "libscipy_openblas*.so")))
    return hits[0] if hits else None


def use_openblas(threads: int) -> dict:
    """Route the oracle's GEMMs to OpenBLAS dgemm/zgemm with `threads` threads."""
    global _blas
    path = openblas_path()
    if path is None:
        raise RuntimeError("scipy's OpenBLAS was not found")
    if _blas is None:
        _blas = C.CDLL(path)
    _blas.scipy_openblas_set_num_threads.argtypes = [C.c_int]
    _blas.scipy_openblas_set_num_threads(int(threads))
    dg = C.cast(_blas.scipy_cblas_dgemm, C.c_void_p)
    zg = C.cast(_blas.scipy_cblas_zgemm, C.c_void_p)
    lib().ora_set_blas(dg, zg)
    lib().ora_set_threads(int(threads))
    cfg = ""
    try:
        _blas.scipy_openblas_get_config.restype = C.c_char_p
        cfg = _blas.scipy_openblas_get_config().decode()
    except Exception:
        pass
    return {"library": os.path.basename(path), "config": cfg, "threads": int(threads)}


def use_naive_gemm() -> None:
    lib().ora_set_blas(None, None)
    lib().ora_set_threads(1)


def _desc(t) -> OraBsdt:
    d = OraBsdt()
    d.rank = t.rank
    d.n_blks = t.n_blks
    d.blk_shape = t.blk_shape.ctypes.data_as(_i64p)
    d.blk_coords = t.coords.ctypes.data_as(_i64p)
    d.blk_dims = t.dims.ctypes.data_as(_i64p)
    d.blk_offsets = t.offsets.ctypes.data_as(_i64p)
    return d


class OraclePlan:
    """DataBlkGenForTenCtrct restated on the CPU."""

    def __init__(self, a, b, a_ctrct, b_ctrct, a_saved=None, b_saved=None):
        if a_saved is None:
            a_saved = [i for i in range(a.rank) if i not in a_ctrct]
        if b_saved is None:
            b_saved = [i for i in range(b.rank) if i not in b_ctrct]
        self.a, self.b = a, b
        self.c_indexes = [a.indexes[i] for i in a_saved] + [b.indexes[i] for i in b_saved]
        self.c_rank = len(a_saved) + len(b_saved)
        ax = [np.ascontiguousarray(np.asarray(v, dtype=np.int32)) for v in (a_ctrct, b_ctrct, a_saved, b_saved)]
        da, db = _desc(a), _desc(b)
        p = lib().ora_plan_build(C.byref(da), C.byref(db), len(a_ctrct), ax[0].ctypes.data_as(_i32p),
                                 ax[1].ctypes.data_as(_i32p), len(a_saved), ax[2].ctypes.data_as(_i32p),
                                 len(b_saved), ax[3].ctypes.data_as(_i32p))
        if not p:
            raise ValueError("oracle: invalid contraction arguments")
        self.handle = C.c_void_p(p)
        self.n_tasks = lib().ora_plan_n_tasks(self.handle)
        self.n_c = lib().ora_plan_n_c(self.handle)
        self.c_raw_size = lib().ora_plan_c_raw_size(self.handle)

    def tasks(self) -> np.ndarray:
        from tensortoolkit_amd.capi import TASK_DTYPE
        ints = np.zeros((self.n_tasks, 9), np.int64)
        fb = np.zeros((self.n_tasks, 2), np.float64)
        if self.n_tasks:
            lib().ora_plan_get_tasks(self.handle, ints.ctypes.data_as(_i64p),
                                     fb.ctypes.data_as(C.POINTER(C.c_double)))
        out = np.zeros(self.n_tasks, dtype=TASK_DTYPE)
        for j, name in enumerate(TASK_DTYPE.names[:9]):
            out[name] = ints[:, j]
        out["f_ex_sign"] = fb[:, 0]
        out["beta"] = fb[:, 1]
        return out

    def c_blocks(self):
        n, r = self.n_c, self.c_rank
        idx = np.zeros(n, np.int64)
        coords = np.zeros((n, r), np.int64)
        dims = np.zeros((n, r), np.int64)
        offs = np.zeros(n, np.int64)
        if n:
            lib().ora_plan_get_c_blocks(self.handle, idx.ctypes.data_as(_i64p), coords.ctypes.data_as(_i64p),
                                        dims.ctypes.data_as(_i64p), offs.ctypes.data_as(_i64p))
        return idx, coords, dims, offs

    def execute(self, sorted_tasks: bool = True) -> np.ndarray:
        """C raw data computed by the restated executor (float64 or complex128)."""
        cplx = self.a.dtype == "c128"
        dt = np.complex128 if cplx else np.float64
        a = np.ascontiguousarray(self.a.data, dtype=dt)
        b = np.ascontiguousarray(self.b.data, dtype=dt)
        c = np.full(max(1, self.c_raw_size), np.nan, dtype=dt)   # never read where beta == 0
        lib().ora_execute(self.handle, int(cplx), a.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p),
                          c.ctypes.data_as(C.c_void_p), int(sorted_tasks))
        return c[:self.c_raw_size]

    def output_tensor(self, dtype):
        from tensortoolkit_amd.blocksparse import BlockSparseTensor
        _, coords, dims, offs = self.c_blocks()
        return BlockSparseTensor(self.c_indexes, coords, dims, dtype=dtype, offsets=offs)

    def close(self):
        if getattr(self, "handle", None) is not None:
            lib().ora_plan_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def part_hash(coords, axes) -> int:
    c = np.ascontiguousarray(np.asarray(coords, dtype=np.int64))
    a = np.ascontiguousarray(np.asarray(axes, dtype=np.int32))
    return int(lib().ora_part_hash(c.ctypes.data_as(_i64p), a.ctypes.data_as(_i32p), len(a)))


def fill_uniform(n: int, seed: int, first: int = 0) -> np.ndarray:
    x = np.empty(n, np.float64)
    lib().ora_fill_uniform(x.ctypes.data_as(C.POINTER(C.c_double)), n, seed, first)
    return x


# ---- brute force and dense helpers (independent of oracle.c) -------------

def brute_force_plan(a, b, a_ctrct, b_ctrct, a_saved=None, b_saved=None):
    """Enumerate every (A block, B block) pair and compare contracted
    coordinates directly -- no hashing -- returning the task list in A-major /
    B-minor order and the set of output block coordinates."""
    if a_saved is None:
        a_saved = [i for i in range(a.rank) if i not in a_ctrct]
    if b_saved is None:
        b_saved = [i for i in range(b.rank) if i not in b_ctrct]
    tasks = []
    for i in range(a.n_blks):
        for j in range(b.n_blks):
            if all(a.coords[i, x] == b.coords[j, y] for x, y in zip(a_ctrct, b_ctrct)):
                c_coords = tuple(int(a.coords[i, x]) for x in a_saved) + tuple(int(b.coords[j, y]) for y in b_saved)
                m = int(np.prod([a.dims[i, x] for x in a_saved])) if a_saved else 1
                k = int(np.prod([a.dims[i, x] for x in a_ctrct])) if a_ctrct else 1
                n = int(np.prod([b.dims[j, y] for y in b_saved])) if b_saved else 1
                tasks.append((i, j, c_coords, m, k, n))
                if not a_saved and not b_saved:
                    break
    return tasks


def dense_contract(a, b, a_ctrct, b_ctrct):
    """np.tensordot of the densified tensors (the reference tests' dense check)."""
    return np.tensordot(a.to_dense(), b.to_dense(), axes=(list(a_ctrct), list(b_ctrct)))
