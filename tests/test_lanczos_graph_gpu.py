"""A Lanczos iteration that lives in one hipGraph (tensortoolkit_amd.lanczos.
lanczos_graph): the basis arena, the Gram-Schmidt coefficients and the
Lanczos coefficients stay in device memory, one iteration is captured once
and replayed.  Checked: replays equal the same launches run eagerly bit for
bit (coefficients, every basis vector, the Ritz vector), with eager work and
synchronisations between the replays; the lowest Ritz value equals the CPU
FP64 Lanczos (oracle/lanczos.py) within 1e-12 (native matvec) or 1e-11
(split matvecs)."""
import ctypes as C

import numpy as np
import pytest

from oracle import lanczos as OL
from tensortoolkit_amd import capi, lanczos as LZ, workloads as W
from tensortoolkit_amd.blocksparse import ContractionPlan
from tests.test_split_gpu import forced_split  # noqa: F401

pytestmark = pytest.mark.gpu


@pytest.fixture
def stream(qlt):
    s = C.c_void_p()
    capi.check(capi.lib().qlt_stream_create(C.byref(s)))
    yield s
    capi.check(capi.lib().qlt_stream_sync(s))
    capi.check(capi.lib().qlt_stream_destroy(s))


def _arena_state(ar):
    capi.check(capi.lib().qlt_stream_sync(ar.hv.stream))
    return (ar.alphas.to_numpy(np.float64), ar.betas.to_numpy(np.float64), ar.V.to_numpy(np.float64),
            ar.count.to_numpy(np.int32))


def _run_replays_with_eager_work(ar, psi0, steps, stream):
    """Capture one iteration, replay it `steps` times; between replays an
    unrelated eager contraction and a device synchronisation."""
    lib = capi.lib()
    mv = ar.hv.mv
    side = ContractionPlan(mv.W, mv.W, [1, 3], [0, 2], stream=stream)
    da = capi.DeviceBuffer.from_numpy(mv.W.data)
    dc = capi.DeviceBuffer(max(1, side.info.c_raw_size) * ar.hv.esz)
    ar.reset(psi0)
    capi.check(lib.qlt_graph_begin(stream))
    ar.iteration()
    g = C.c_void_p()
    capi.check(lib.qlt_graph_end(stream, C.byref(g)))
    for _ in range(steps):
        capi.check(lib.qlt_graph_replay(g, stream))
        side.execute(mv.dtype, capi.QLT_MODE_NATIVE, da.ptr, da.ptr, dc.ptr, stream=stream)
        capi.check(lib.qlt_device_sync())
    capi.check(lib.qlt_graph_destroy(g))
    return _arena_state(ar)


@pytest.mark.parametrize("u1u1,dtype,D,mode", [
    (False, "f64", 96, capi.QLT_MODE_NATIVE),
    (True, "c128", 64, capi.QLT_MODE_NATIVE),
    (True, "c128", 64, capi.QLT_MODE_SPLIT_BF16),
    (True, "c128", 64, capi.QLT_MODE_SPLIT_I8),
], ids=["u1-f64-native", "u1u1-c128-native", "u1u1-c128-bf16", "u1u1-c128-i8"])
def test_graph_iteration_matches_eager_and_cpu(forced_split, stream, u1u1, dtype, D, mode):
    steps = 12
    print(f"case: lanczos graph u1u1={u1u1} {dtype} D={D} mode={mode}", flush=True)
    mv = W.hermitian_matvec(W.small_matvec(D=D, dtype=dtype, u1u1=u1u1))
    hv = LZ.DeviceMatvec(mv, mode, stream)
    ar = LZ.KrylovArena(hv, steps)
    psi0 = mv.psi.data.copy()
    # eager launches of the same iterations
    ar.reset(psi0)
    for _ in range(steps):
        ar.iteration()
    eager = _arena_state(ar)
    assert eager[3][0] == steps + 1
    # one captured iteration, replayed, with eager work in between
    replay = _run_replays_with_eager_work(ar, psi0, steps, stream)
    for e, r, name in zip(eager, replay, ("alphas", "betas", "basis", "count")):
        np.testing.assert_array_equal(r, e, err_msg=name)
    # the driver, graph and eager, against the CPU FP64 Lanczos
    res_g = LZ.lanczos_graph(hv, psi0, steps, use_graph=True, arena=ar)
    gs_g = res_g.ground_state.to_numpy(mv.psi.np_dtype, mv.psi.raw_size)
    res_e = LZ.lanczos_graph(hv, psi0, steps, use_graph=False, arena=ar)
    gs_e = res_e.ground_state.to_numpy(mv.psi.np_dtype, mv.psi.raw_size)
    np.testing.assert_array_equal(gs_g, gs_e)
    assert res_g.eigenvalue == res_e.eigenvalue
    theta, gs_ref, *_ = OL.lanczos(mv, psi0, steps)
    tol = 1e-12 if mode == capi.QLT_MODE_NATIVE else 1e-11
    assert abs(res_g.eigenvalue - theta) <= tol * abs(theta)
    assert abs(abs(np.vdot(gs_ref, gs_g)) - 1.0) <= 1e-9
    # the basis is orthonormal to working precision
    V = replay[2].view(mv.psi.np_dtype).reshape(steps + 1, -1)
    G = V.conj() @ V.T
    assert np.max(np.abs(G - np.eye(steps + 1))) <= 1e-12
