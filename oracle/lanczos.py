"""CPU reference of the end-to-end two-site update -- TEST INFRASTRUCTURE ONLY.

The same Lanczos iteration as tensortoolkit_amd/lanczos.py (full twice-
repeated Gram-Schmidt reorthogonalisation, FP64 tridiagonal eigensolve) in
numpy, with every matvec computed by the oracle's FP64 contraction (the
reference algorithm restated in C, BLAS per task).  Only ``tests/`` and the
end-to-end mode of ``bench.py`` use it, as the full-FP64 CPU run the device
result is compared with.
"""
from __future__ import annotations

import time

import numpy as np

from tensortoolkit_amd import workloads as W
from . import oracle as O


class OracleMatvec:
    def __init__(self, mv: W.Matvec):
        self.plans = []
        cur = mv.psi
        for step, axes in enumerate(W.MATVEC_AXES):
            op = (mv.L, mv.W, mv.W, mv.R)[step]
            first, second = (op, cur) if step == 0 else (cur, op)
            p = O.OraclePlan(first, second, list(axes[0]), list(axes[1]))
            self.plans.append(p)
            cur = p.output_tensor(mv.dtype)

    def __call__(self, x: np.ndarray) -> np.ndarray:
        self.plans[0].b.data = x
        y = self.plans[0].execute()
        for p in self.plans[1:]:
            p.a.data = y
            y = p.execute()
        return y


def lanczos(mv: W.Matvec, psi0: np.ndarray, steps: int = 20):
    """Returns (lowest Ritz value, normalised Ritz vector, alphas, betas, seconds)."""
    H = OracleMatvec(mv)
    t0 = time.perf_counter()
    v = psi0 / np.sqrt(np.vdot(psi0, psi0).real)
    basis = [v]
    alphas, betas = [], []
    for j in range(steps):
        w = H(basis[j])
        a = np.vdot(basis[j], w).real
        alphas.append(a)
        w = w - a * basis[j]
        if j > 0:
            w = w - betas[-1] * basis[j - 1]
        for _ in range(2):
            for u in basis:
                w = w - np.vdot(u, w) * u
        b = np.sqrt(np.vdot(w, w).real)
        if j == steps - 1 or b == 0.0:
            break
        betas.append(b)
        basis.append(w / b)
    k = len(alphas)
    T = np.diag(alphas) + np.diag(betas[:k - 1], 1) + np.diag(betas[:k - 1], -1)
    evals, evecs = np.linalg.eigh(T)
    gs = sum(evecs[i, 0] * basis[i] for i in range(k))
    gs = gs / np.sqrt(np.vdot(gs, gs).real)
    return float(evals[0]), gs, np.array(alphas), np.array(betas), time.perf_counter() - t0
