"""Writes tests/golden/e2e_config4.npz (run from the repository root:
``python tests/golden/make_e2e_golden.py``): the full-FP64 CPU run of the
end-to-end two-site update on the U(1)xU(1) D = 4096 complex128 tensors made
Hermitian (tensortoolkit_amd.workloads.config4, fixed seed), i.e. 20 Lanczos
steps with the oracle's FP64 matvec (the reference algorithm restated in C,
OpenBLAS per task) and full reorthogonalisation, then the oracle's block SVD
of the Ritz vector over (l, s1 | s2, r), truncated at cutoff 1e-10 with
Dmax = 4096 (ten_svd.h:185-372 semantics).  Stored: the lowest Ritz value,
the Lanczos coefficients, the truncation error, the kept dimension D and the
kept singular values, as the expected outputs of the device run
(tests/test_e2e_gpu.py)."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import decomp as OD, lanczos as OL, oracle as O  # noqa: E402
from tensortoolkit_amd import workloads as W  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "e2e_config4.npz")
STEPS, CUTOFF, DMAX = 20, 1e-10, 4096


def main():
    O.use_openblas(os.cpu_count() or 1)
    mv = W.config4()
    theta, gs, alphas, betas, secs = OL.lanczos(mv, mv.psi.data.copy(), STEPS)
    psi = mv.psi
    t = type(psi)(psi.indexes, psi.coords, psi.dims, dtype=psi.dtype, data=gs, offsets=psi.offsets)
    out, err, D, svals = OD.svd(t, 2, CUTOFF, 1, DMAX)
    kept = np.sort(np.concatenate([s for _, _, s, *_ in out]))[::-1]
    np.savez_compressed(OUT, eigenvalue=np.array([theta]), alphas=alphas, betas=betas,
                        trunc_err=np.array([err]), D=np.array([D]), kept_svals=kept,
                        steps=np.array([STEPS]), cutoff=np.array([CUTOFF]), dmax=np.array([DMAX]))
    print(f"{mv.name}: eigenvalue {theta!r}, trunc err {err!r}, D {D}, lanczos {secs:.1f} s")


if __name__ == "__main__":
    main()
