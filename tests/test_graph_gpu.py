"""Graph capture of the 4-contraction matvec chain (qlt_graph_begin / end /
replay), with the tile classes forked onto the library's class streams inside
the capture.  Every replay -- with eager work and a changed input between
replays -- must equal the eager chain on the same input bit for bit (the
kernels, their order and every reduction order are the same)."""
import ctypes as C

import numpy as np
import pytest

from tensortoolkit_amd import capi, workloads as W
from tensortoolkit_amd.blocksparse import ContractionPlan
from tests.test_split_gpu import forced_split  # noqa: F401

pytestmark = pytest.mark.gpu


def _chain(mv, mode, stream):
    esz = 16 if mv.dtype == "c128" else 8
    bufs = {k: capi.DeviceBuffer.from_numpy(getattr(mv, k).data) for k in ("L", "psi", "W", "R")}
    plans, outs = [], []
    x = mv.psi
    for step, axes in enumerate(W.MATVEC_AXES):
        op = (mv.L, mv.W, mv.W, mv.R)[step]
        first, second = (op, x) if step == 0 else (x, op)
        p = ContractionPlan(first, second, list(axes[0]), list(axes[1]), stream=stream)
        p.prepare(mv.dtype, mode)
        o = p.output_tensor(mv.dtype)
        plans.append(p)
        outs.append(capi.DeviceBuffer(max(1, o.raw_size) * esz))
        x = o

    def run():
        args = [(bufs["L"].ptr, bufs["psi"].ptr), (outs[0].ptr, bufs["W"].ptr), (outs[1].ptr, bufs["W"].ptr),
                (outs[2].ptr, bufs["R"].ptr)]
        for s, (a, b) in enumerate(args):
            plans[s].execute(mv.dtype, mode, a, b, outs[s].ptr, stream=stream)

    return plans, bufs, outs, run, x


def _replays(mode, mv=None):
    lib = capi.lib()
    stream = C.c_void_p()
    capi.check(lib.qlt_stream_create(C.byref(stream)))
    if mv is None:
        mv = W.small_matvec(D=64, dtype="c128", u1u1=True)
    assert mv.dtype == "c128"
    plans, bufs, outs, run, out_t = _chain(mv, mode, stream)
    npd = np.complex128

    def result():
        capi.check(lib.qlt_stream_sync(stream))
        return outs[-1].to_numpy(npd, out_t.raw_size)

    # eager reference for two different inputs
    run()
    eager0 = result()
    scale = (C.c_double * 2)(0.5, -0.25)
    zero = (C.c_double * 2)(0.0, 0.0)
    psi2 = bufs["psi"].to_numpy(npd, mv.psi.raw_size) * complex(0.5, -0.25)
    capi.check(lib.qlt_graph_begin(stream))
    run()
    g = C.c_void_p()
    capi.check(lib.qlt_graph_end(stream, C.byref(g)))
    try:
        for rep in range(3):
            capi.check(lib.qlt_graph_replay(g, stream))
            assert np.array_equal(result(), eager0), f"replay {rep} differs from the eager chain"
            run()                                    # eager work between replays
            assert np.array_equal(result(), eager0)
        # change the input in place (eager), replay, compare with the eager chain on it
        capi.check(lib.qlt_axpby(capi.QLT_C128, mv.psi.raw_size, scale, bufs["psi"].ptr, zero, bufs["psi"].ptr,
                                 stream))
        capi.check(lib.qlt_stream_sync(stream))
        assert np.allclose(bufs["psi"].to_numpy(npd, mv.psi.raw_size), psi2, rtol=1e-15, atol=1e-300)
        capi.check(lib.qlt_graph_replay(g, stream))
        rep_out = result()
        run()
        eager1 = result()
        assert np.array_equal(rep_out, eager1)
        assert not np.array_equal(eager1, eager0)
        for rep in range(2):
            capi.check(lib.qlt_graph_replay(g, stream))
            assert np.array_equal(result(), eager1)
    finally:
        capi.check(lib.qlt_graph_destroy(g))
        capi.check(lib.qlt_stream_destroy(stream))


def test_graph_replay_native(qlt):
    print("case: graph replay, native, small U(1)xU(1) matvec", flush=True)
    _replays(capi.QLT_MODE_NATIVE)


@pytest.mark.parametrize("mode", [capi.QLT_MODE_SPLIT_BF16, capi.QLT_MODE_SPLIT_I8], ids=["bf16", "i8"])
def test_graph_replay_split(qlt, forced_split, mode):
    print(f"case: graph replay, split mode {mode} forced, small U(1)xU(1) matvec", flush=True)
    _replays(mode)


@pytest.mark.parametrize("mode", [capi.QLT_MODE_NATIVE, capi.QLT_MODE_SPLIT_BF16, capi.QLT_MODE_SPLIT_I8],
                         ids=["native", "bf16", "i8"])
def test_graph_replay_full_size(qlt, mode):
    """The D = 4096 U(1)xU(1) complex128 chain at the default dispatch: k-sliced
    tile classes with their partial-sum reductions, the split GEMMs on L.psi
    and every class stream forked inside the capture."""
    print(f"case: graph replay, mode {mode}, D = 4096 U(1)xU(1) complex128 matvec", flush=True)
    _replays(mode, W.config3())
