"""GPU Lanczos on the device matvec (tensortoolkit_amd.lanczos) against the
same iteration in FP64 on the CPU (oracle/lanczos.py): lowest Ritz value
within 1e-12 relative (native FP64 matvec) or 1e-11 (split-BF16 matvec),
Ritz vectors with overlap 1 - 1e-9."""
import numpy as np
import pytest

from oracle import lanczos as OL
from tensortoolkit_amd import capi, lanczos as LZ, workloads as W
from tests.test_split_gpu import forced_split  # noqa: F401

pytestmark = pytest.mark.gpu


def _run(mv, mode, steps):
    hv = LZ.DeviceMatvec(mv, mode)
    res = LZ.lanczos(hv, mv.psi.data.copy(), steps)
    gs = res.ground_state.to_numpy(mv.psi.np_dtype, mv.psi.raw_size)
    return res, gs


@pytest.mark.parametrize("u1u1,dtype,D", [(False, "f64", 96), (True, "c128", 64)])
def test_lanczos_native_matches_cpu(qlt, u1u1, dtype, D):
    print(f"case: lanczos native u1u1={u1u1} D={D}", flush=True)
    mv = W.hermitian_matvec(W.small_matvec(D=D, dtype=dtype, u1u1=u1u1))
    res, gs = _run(mv, capi.QLT_MODE_NATIVE, 12)
    theta, gs_ref, *_ = OL.lanczos(mv, mv.psi.data.copy(), 12)
    assert abs(res.eigenvalue - theta) <= 1e-12 * abs(theta)
    assert abs(abs(np.vdot(gs_ref, gs)) - 1.0) <= 1e-9


@pytest.mark.parametrize("mode", [capi.QLT_MODE_SPLIT_BF16, capi.QLT_MODE_SPLIT_I8], ids=["bf16", "i8"])
def test_lanczos_split_matches_cpu(qlt, forced_split, mode):
    print(f"case: lanczos split mode {mode} u1u1 D=64", flush=True)
    mv = W.hermitian_matvec(W.small_matvec(D=64, dtype="c128", u1u1=True))
    res, gs = _run(mv, mode, 12)
    theta, gs_ref, *_ = OL.lanczos(mv, mv.psi.data.copy(), 12)
    assert abs(res.eigenvalue - theta) <= 1e-11 * abs(theta)
    assert abs(abs(np.vdot(gs_ref, gs)) - 1.0) <= 1e-9
