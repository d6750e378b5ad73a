"""GPU: the fused Gram-Schmidt step qlt_basis_update_dot (w += scale V c, then
c_out = V^H w) against qlt_basis_update followed by qlt_basis_dot: the same w
and the same c_out bit for bit, with m active vectors on both sides of the
one-pass limit (m <= 16); both against numpy in FP64.  The primitives stand
for LinearCombine (ten_linear_combine.h:186) and InnerProduct
(basic_operations.h:146-203) of the reference's Lanczos solver."""
import ctypes as C

import numpy as np
import pytest

from tensortoolkit_amd import capi

pytestmark = pytest.mark.gpu


def _dev(a):
    return capi.DeviceBuffer.from_numpy(np.ascontiguousarray(a))


@pytest.mark.parametrize("dtype", ["f64", "c128"])
@pytest.mark.parametrize("n,m", [(1000, 1), (1000, 16), (100003, 5), (100003, 16), (100003, 17), (4099, 24)])
def test_basis_update_dot_matches_update_then_dot(qlt, dtype, n, m):
    print(f"case: basis update+dot {dtype} n={n} m={m}", flush=True)
    lib = capi.lib()
    k_max = 24
    code = capi.QLT_C128 if dtype == "c128" else capi.QLT_F64
    rng = np.random.default_rng(1000 * m + n % 97)
    if dtype == "c128":
        V = rng.uniform(-1, 1, (k_max, n)) + 1j * rng.uniform(-1, 1, (k_max, n))
        w = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
        c = rng.uniform(-1, 1, k_max) + 1j * rng.uniform(-1, 1, k_max)
    else:
        V = rng.uniform(-1, 1, (k_max, n))
        w = rng.uniform(-1, 1, n)
        c = rng.uniform(-1, 1, k_max) + 0j
    scale = -0.75
    c_pairs = np.stack([c.real, c.imag], axis=1).reshape(-1).copy()
    dV, dc, dcount = _dev(V), _dev(c_pairs), _dev(np.array([m], dtype=np.int32))
    work_n = 2 * 512 * k_max + 1
    # two calls
    dw1, dout1, dwork1 = _dev(w), capi.DeviceBuffer(16 * k_max), capi.DeviceBuffer(8 * work_n)
    capi.check(lib.qlt_basis_update(code, n, k_max, dcount.ptr, scale, dc.ptr, dV.ptr, n, dw1.ptr, None))
    capi.check(lib.qlt_basis_dot(code, n, k_max, dcount.ptr, dV.ptr, n, dw1.ptr, dout1.ptr, dwork1.ptr, None))
    # the fused call
    dw2, dout2, dwork2 = _dev(w), capi.DeviceBuffer(16 * k_max), capi.DeviceBuffer(8 * work_n)
    capi.check(lib.qlt_basis_update_dot(code, n, k_max, dcount.ptr, scale, dc.ptr, dV.ptr, n, dw2.ptr, dout2.ptr,
                                        dwork2.ptr, None))
    capi.check(lib.qlt_device_sync())
    npt = np.complex128 if dtype == "c128" else np.float64
    w1, w2 = dw1.to_numpy(npt, n), dw2.to_numpy(npt, n)
    o1, o2 = dout1.to_numpy(np.float64, 2 * k_max), dout2.to_numpy(np.float64, 2 * k_max)
    np.testing.assert_array_equal(w2, w1)
    np.testing.assert_array_equal(o2, o1)
    # FP64 reference
    cm = c[:m] if dtype == "c128" else c[:m].real
    w_ref = w + scale * (cm @ V[:m])
    o_ref = V[:m].conj() @ w_ref
    assert np.linalg.norm(w2 - w_ref) <= 1e-13 * np.linalg.norm(w_ref)
    got = o2[0::2] + 1j * o2[1::2]
    assert np.linalg.norm(got[:m] - o_ref) <= 1e-12 * np.linalg.norm(o_ref)
    assert np.all(got[m:] == 0)
    # c and c_out may not alias
    st = lib.qlt_basis_update_dot(code, n, k_max, dcount.ptr, scale, dc.ptr, dV.ptr, n, dw2.ptr, dc.ptr, dwork2.ptr,
                                  None)
    assert st == 1   # QLT_ERR_INVALID_ARG
