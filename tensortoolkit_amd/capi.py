"""ctypes binding of the C ABI in ``include/qlten_amd.h``.

This is plumbing for the Python tests and the benchmark; C++ callers use the
``qlten`` headers in ``include/qlten/``.  Loading fails loudly when the HIP
library has not been built: there is no CPU fallback anywhere in the product.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libqlten_amd.so"

QLT_OK = 0
QLT_F64, QLT_C128, QLT_F32, QLT_C64 = 0, 1, 2, 3
QLT_MODE_NATIVE, QLT_MODE_SPLIT_BF16, QLT_MODE_SPLIT_I8 = 0, 1, 2
QLT_MAX_RANK = 8

_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


class QltError(RuntimeError):
    """A libqlten_amd call returned a non-zero status."""

    def __init__(self, status: int, message: str):
        super().__init__(f"qlt status {status}: {message}")
        self.status = status


class BsdtDesc(C.Structure):
    _fields_ = [
        ("rank", C.c_int32),
        ("n_blks", C.c_int64),
        ("blk_shape", _i64p),
        ("blk_coords", _i64p),
        ("blk_dims", _i64p),
        ("blk_offsets", _i64p),
    ]


QLT_TUNE_BUCKETS = 16
QLT_TILE_CLASSES = 6


class Tuning(C.Structure):
    """qlt_tuning: tile class per size bucket (-1: built-in rule) and the
    k-slicing workgroup target per tile class (0: built-in, -1: unsliced)."""
    _fields_ = [
        ("tile_class", (C.c_int8 * QLT_TUNE_BUCKETS) * QLT_TUNE_BUCKETS),
        ("slice_target", C.c_int32 * QLT_TILE_CLASSES),
    ]


class TuneItem(C.Structure):
    _fields_ = [
        ("a", C.POINTER(BsdtDesc)),
        ("b", C.POINTER(BsdtDesc)),
        ("n_ctrct", C.c_int32),
        ("a_ctrct_axes", _i32p),
        ("b_ctrct_axes", _i32p),
        ("n_a_saved", C.c_int32),
        ("a_saved_axes", _i32p),
        ("n_b_saved", C.c_int32),
        ("b_saved_axes", _i32p),
        ("a_data", C.c_void_p),
        ("b_data", C.c_void_p),
        ("c_data", C.c_void_p),
    ]


class Task(C.Structure):
    _fields_ = [(name, C.c_int64) for name in (
        "a_blk_idx", "a_data_offset", "b_blk_idx", "b_data_offset",
        "c_blk_idx", "c_data_offset", "m", "k", "n")] + [
        ("f_ex_sign", C.c_double), ("beta", C.c_double)]


class PlanInfo(C.Structure):
    _fields_ = [
        ("n_tasks", C.c_int64),
        ("n_c_blks", C.c_int64),
        ("c_raw_size", C.c_int64),
        ("c_rank", C.c_int32),
        ("c_is_scalar", C.c_int32),
        ("sum_mkn", C.c_int64),
        ("n_tiles", C.c_int64),
        ("a_unique_elems", C.c_int64),
        ("b_unique_elems", C.c_int64),
        ("n_panel_tiles", C.c_int64),
    ]


TASK_DTYPE = np.dtype([(name, np.int64) for name in (
    "a_blk_idx", "a_data_offset", "b_blk_idx", "b_data_offset",
    "c_blk_idx", "c_data_offset", "m", "k", "n")] + [
    ("f_ex_sign", np.float64), ("beta", np.float64)])
assert TASK_DTYPE.itemsize == C.sizeof(Task)

_lib = None

_SIGS = {
    "qlt_last_error": ([], C.c_char_p),
    "qlt_init": ([C.c_int32], C.c_int32),
    "qlt_finalize": ([], C.c_int32),
    "qlt_device_sync": ([], C.c_int32),
    "qlt_stream_create": ([C.POINTER(C.c_void_p)], C.c_int32),
    "qlt_stream_destroy": ([C.c_void_p], C.c_int32),
    "qlt_stream_sync": ([C.c_void_p], C.c_int32),
    "qlt_malloc": ([C.POINTER(C.c_void_p), C.c_size_t], C.c_int32),
    "qlt_free": ([C.c_void_p], C.c_int32),
    "qlt_memcpy_h2d": ([C.c_void_p, C.c_void_p, C.c_size_t], C.c_int32),
    "qlt_memcpy_d2h": ([C.c_void_p, C.c_void_p, C.c_size_t], C.c_int32),
    "qlt_memcpy_d2d": ([C.c_void_p, C.c_void_p, C.c_size_t], C.c_int32),
    "qlt_memset_zero": ([C.c_void_p, C.c_size_t, C.c_void_p], C.c_int32),
    "qlt_plan_build": ([C.POINTER(BsdtDesc), C.POINTER(BsdtDesc), C.c_int32, _i32p, _i32p,
                        C.c_int32, _i32p, C.c_int32, _i32p, C.c_void_p, C.POINTER(C.c_void_p)],
                       C.c_int32),
    "qlt_plan_destroy": ([C.c_void_p], C.c_int32),
    "qlt_plan_get_info": ([C.c_void_p, C.POINTER(PlanInfo)], C.c_int32),
    "qlt_plan_get_tasks": ([C.c_void_p, C.c_void_p], C.c_int32),
    "qlt_plan_get_c_blocks": ([C.c_void_p, _i64p, _i64p, _i64p, _i64p], C.c_int32),
    "qlt_plan_set_c_offsets": ([C.c_void_p, _i64p], C.c_int32),
    "qlt_plan_prepare": ([C.c_void_p, C.c_int32, C.c_int32], C.c_int32),
    "qlt_contract_execute": ([C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                              _f64p, _f64p, C.c_void_p], C.c_int32),
    "qlt_set_split_min_mean_k": ([C.c_double], C.c_int32),
    "qlt_tuning_get": ([C.POINTER(Tuning)], C.c_int32),
    "qlt_tuning_set": ([C.POINTER(Tuning)], C.c_int32),
    "qlt_autotune": ([C.c_int32, C.POINTER(TuneItem), C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                      C.POINTER(Tuning), C.POINTER(C.c_double), C.POINTER(C.c_double), _i32p], C.c_int32),
    "qlt_get_split_min_mean_k": ([], C.c_double),
    "qlt_set_row_panels": ([C.c_int32], C.c_int32),
    "qlt_get_row_panels": ([], C.c_int32),
    "qlt_kernel_probe_set": ([C.c_void_p, C.c_void_p], C.c_int32),
    "qlt_kernel_probe_last": ([C.c_char_p, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)], C.c_int32),
    "qlt_graph_begin": ([C.c_void_p], C.c_int32),
    "qlt_graph_end": ([C.c_void_p, C.POINTER(C.c_void_p)], C.c_int32),
    "qlt_graph_replay": ([C.c_void_p, C.c_void_p], C.c_int32),
    "qlt_graph_destroy": ([C.c_void_p], C.c_int32),
    "qlt_fill_uniform": ([C.c_void_p, C.c_int64, C.c_uint64, C.c_int64, C.c_double, C.c_double, C.c_void_p],
                         C.c_int32),
    "qlt_axpby": ([C.c_int32, C.c_int64, _f64p, C.c_void_p, _f64p, C.c_void_p, C.c_void_p], C.c_int32),
    "qlt_dot": ([C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, _f64p, C.c_void_p], C.c_int32),
    "qlt_real_to_complex": ([C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int32),
    "qlt_conjugate": ([C.c_int64, C.c_void_p, C.c_void_p], C.c_int32),
    "qlt_scale_segments": ([C.c_int32, C.c_int64, _i64p, _i64p, _f64p, C.c_void_p, C.c_void_p], C.c_int32),
    "qlt_scale_segments_noalloc": ([C.c_int32, C.c_int64, _i64p, _i64p, _f64p, C.c_void_p, C.c_void_p],
                                   C.c_int32),
    "qlt_conjugate_c64": ([C.c_int64, C.c_void_p, C.c_void_p], C.c_int32),
    "qlt_fill_uniform_f32": ([C.c_void_p, C.c_int64, C.c_uint64, C.c_int64, C.c_double, C.c_double, C.c_void_p],
                             C.c_int32),
    "qlt_convert": ([C.c_int32, C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int32),
    "qlt_basis_dot": ([C.c_int32, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                       C.c_void_p, C.c_void_p], C.c_int32),
    "qlt_basis_update": ([C.c_int32, C.c_int64, C.c_int32, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p,
                          C.c_int64, C.c_void_p, C.c_void_p], C.c_int32),
    "qlt_basis_update_dot": ([C.c_int32, C.c_int64, C.c_int32, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p,
                              C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int32),
    "qlt_lanczos_extend": ([C.c_int32, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                            C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int32),
    "qlt_basis_select": ([C.c_int32, C.c_int64, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_int64,
                          C.c_void_p, C.c_void_p], C.c_int32),
    "qlt_copy_blocks_2d": ([C.c_int32, C.c_int64, _i64p, _i64p, C.c_void_p, _i64p, _i64p, C.c_void_p, _i64p, _i64p,
                            C.c_void_p], C.c_int32),
    "qlt_block_svd": ([C.c_int32, C.c_int64, _i64p, _i64p, C.c_void_p, _i64p, C.c_void_p, _i64p, C.c_void_p, _i64p,
                       C.c_void_p, _i64p, C.POINTER(C.c_int32), C.c_void_p], C.c_int32),
    "qlt_block_qr": ([C.c_int32, C.c_int64, _i64p, _i64p, C.c_void_p, _i64p, C.c_void_p, _i64p, C.c_void_p, _i64p,
                      C.c_void_p], C.c_int32),
}

EXPORTED_SYMBOLS = tuple(_SIGS)


def lib() -> C.CDLL:
    """Load ``libqlten_amd.so`` (RuntimeError if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: run tensortoolkit_amd.build (or __graft_entry__.build()) first")
        handle = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | os.RTLD_GLOBAL)
        for name, (args, res) in _SIGS.items():
            fn = getattr(handle, name)
            fn.argtypes = args
            fn.restype = res
        _lib = handle
    return _lib


def check(status: int) -> None:
    if status != QLT_OK:
        raise QltError(status, lib().qlt_last_error().decode(errors="replace"))


def i64(a) -> C.POINTER(C.c_int64):
    """Pointer to a C-contiguous int64 numpy array."""
    return a.ctypes.data_as(_i64p)


def i32(a) -> C.POINTER(C.c_int32):
    return a.ctypes.data_as(_i32p)


def f64(a) -> C.POINTER(C.c_double):
    return a.ctypes.data_as(_f64p)


_initialized_device = None


def init(device: int = 0) -> None:
    global _initialized_device
    if _initialized_device != device:
        check(lib().qlt_init(device))
        _initialized_device = device


class DeviceBuffer:
    """Device allocation made with qlt_malloc; freed on close()/GC."""

    def __init__(self, nbytes: int):
        init()
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib().qlt_malloc(C.byref(p), self.nbytes))
        self.ptr = p

    @classmethod
    def from_numpy(cls, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes)
        if arr.nbytes:
            check(lib().qlt_memcpy_h2d(buf.ptr, arr.ctypes.data_as(C.c_void_p), arr.nbytes))
        return buf

    def to_numpy(self, dtype, count: int | None = None) -> np.ndarray:
        dtype = np.dtype(dtype)
        if count is None:
            count = self.nbytes // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            check(lib().qlt_memcpy_d2h(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes))
        return out

    def close(self) -> None:
        if getattr(self, "ptr", None) is not None and self.ptr.value is not None:
            lib().qlt_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
