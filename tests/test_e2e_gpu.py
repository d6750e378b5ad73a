"""The end-to-end two-site update at full size on the GPU (U(1)xU(1) D = 4096
complex128 tensors made Hermitian, fixed seed): 20 Lanczos steps with the
matvec in each split mode and FP64 reorthogonalisation on the device
(tensortoolkit_amd.lanczos), then the FP64 batched block SVD of the Ritz
vector over (l, s1 | s2, r) truncated at cutoff 1e-10, Dmax 4096
(tensortoolkit_amd.decompose.svd; ten_svd.h:185-372,
data_blk_operations.h:630-663).  Expected values: the same computation in
full FP64 on the CPU (the oracle's matvec and block SVD), stored in
tests/golden/e2e_config4.npz by tests/golden/make_e2e_golden.py."""
import ctypes as C
import os

import numpy as np
import pytest

from tensortoolkit_amd import capi, decompose as DEC, lanczos as LZ, workloads as W

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "e2e_config4.npz")


@pytest.fixture(scope="module")
def config4():
    return W.config4()


@pytest.mark.parametrize("mode,driver", [(capi.QLT_MODE_SPLIT_BF16, "graph"), (capi.QLT_MODE_SPLIT_I8, "graph"),
                                         (capi.QLT_MODE_NATIVE, "graph"), (capi.QLT_MODE_SPLIT_BF16, "loop")],
                         ids=["split_bf16-graph", "split_i8-graph", "native-graph", "split_bf16-loop"])
def test_two_site_update_matches_fp64_cpu(qlt, config4, mode, driver):
    g = np.load(GOLDEN)
    steps, cutoff, dmax = int(g["steps"][0]), float(g["cutoff"][0]), int(g["dmax"][0])
    mv = config4
    print(f"case: e2e {mv.name} mode {mode} {driver} lanczos {steps} steps, svd cutoff {cutoff} Dmax {dmax}",
          flush=True)
    stream = C.c_void_p()
    capi.check(capi.lib().qlt_stream_create(C.byref(stream)))
    hv = LZ.DeviceMatvec(mv, mode, stream)
    if driver == "graph":
        lz = LZ.lanczos_graph(hv, mv.psi.data.copy(), steps)
    else:
        lz = LZ.lanczos(hv, mv.psi.data.copy(), steps)
    print(f"lanczos {lz.seconds:.3f} s ({lz.matvec_seconds:.3f} s iterating)", flush=True)
    theta = float(g["eigenvalue"][0])
    rel = abs(lz.eigenvalue - theta) / abs(theta)
    print(f"eigenvalue {lz.eigenvalue!r} vs {theta!r}: rel err {rel:.3e}", flush=True)
    assert rel <= 1e-12
    res = DEC.svd(mv.psi, 2, (0, 0), cutoff, 1, dmax, t_dev=lz.ground_state, stream=stream)
    err_ref, D_ref = float(g["trunc_err"][0]), int(g["D"][0])
    print(f"trunc err {res.trunc_err!r} vs {err_ref!r}, D {res.D} vs {D_ref}, svd sweeps {res.sweeps}, "
          f"{len(res.svals)} block matrices", flush=True)
    assert res.D == D_ref
    assert abs(res.trunc_err - err_ref) <= 1e-12
    kept = np.sort(np.concatenate([np.asarray(s) for s in res.svals]))[::-1][:D_ref]
    # singular values move by at most the change of the (unit) Ritz vector
    assert np.max(np.abs(kept - g["kept_svals"])) <= 1e-9
    capi.check(capi.lib().qlt_stream_sync(stream))
    capi.check(capi.lib().qlt_stream_destroy(stream))
