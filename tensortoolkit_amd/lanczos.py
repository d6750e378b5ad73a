"""Lanczos ground state of the two-site effective Hamiltonian on the device.

The matvec psi' = L . psi . W1 . W2 . R runs as four grouped contractions on
cached device plans (native FP64, split-BF16 or split-int8 mode); the Krylov vectors
stay in HBM and every vector operation is a device kernel in FP64 -- the
reference's LinearCombine (ten_linear_combine.h:186), InnerProduct
(basic_operations.h:203) and Normalize (qltensor_impl.h:568-573) building
blocks (qlt_axpby, qlt_dot) -- with full (twice-repeated classical
Gram-Schmidt) reorthogonalisation.  Only the tridiagonal matrix (alpha, beta)
goes to the host, where its eigenproblem is solved in FP64.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from . import capi, workloads as W
from .blocksparse import ContractionPlan


class DeviceMatvec:
    """Cached plans and buffers of one effective-Hamiltonian matvec."""

    def __init__(self, mv: W.Matvec, mode: int = capi.QLT_MODE_SPLIT_BF16, stream=None):
        self.mv, self.mode, self.stream, self.dtype = mv, mode, stream, mv.dtype
        self.esz = 16 if mv.dtype == "c128" else 8
        self.n = mv.psi.raw_size
        self.ops = {k: capi.DeviceBuffer.from_numpy(getattr(mv, k).data) for k in ("L", "W", "R")}
        self.plans, self.bufs = [], []
        x = mv.psi
        for step, axes in enumerate(W.MATVEC_AXES):
            op = (mv.L, mv.W, mv.W, mv.R)[step]
            first, second = (op, x) if step == 0 else (x, op)
            plan = ContractionPlan(first, second, list(axes[0]), list(axes[1]), stream=stream)
            plan.prepare(self.dtype, mode)
            out = plan.output_tensor(self.dtype)
            self.plans.append(plan)
            if step < 3:
                self.bufs.append(capi.DeviceBuffer(max(1, out.raw_size) * self.esz))
            x = out
        if not (x.n_blks == mv.psi.n_blks and np.array_equal(x.offsets, mv.psi.offsets)):
            raise ValueError("the matvec output does not have psi's block layout")
        self.flops = sum(p.flops(self.dtype) for p in self.plans)

    def apply(self, x_ptr, y_ptr) -> None:
        """y = H x (device pointers, psi's raw layout)."""
        ops = self.ops
        args = [(ops["L"].ptr, x_ptr, self.bufs[0].ptr), (self.bufs[0].ptr, ops["W"].ptr, self.bufs[1].ptr),
                (self.bufs[1].ptr, ops["W"].ptr, self.bufs[2].ptr), (self.bufs[2].ptr, ops["R"].ptr, y_ptr)]
        for plan, (a, b, c) in zip(self.plans, args):
            plan.execute(self.dtype, self.mode, a, b, c, stream=self.stream)


def _cpair(z: complex) -> C.Array:
    return (C.c_double * 2)(float(np.real(z)), float(np.imag(z)))


class DeviceVectors:
    """FP64 vector kernels on psi-shaped device buffers."""

    def __init__(self, dtype: str, n: int, stream=None):
        self.dt = capi.QLT_C128 if dtype == "c128" else capi.QLT_F64
        self.n, self.stream = n, stream
        self.lib = capi.lib()

    def dot(self, x, y) -> complex:
        out = (C.c_double * 2)()
        capi.check(self.lib.qlt_dot(self.dt, self.n, x, y, out, self.stream))
        return complex(out[0], out[1])

    def axpby(self, a: complex, x, b: complex, y) -> None:
        capi.check(self.lib.qlt_axpby(self.dt, self.n, _cpair(a), x, _cpair(b), y, self.stream))


@dataclass
class LanczosResult:
    eigenvalue: float
    alphas: np.ndarray
    betas: np.ndarray
    ground_state: capi.DeviceBuffer
    seconds: float
    matvec_seconds: float


def lanczos(hv: DeviceMatvec, psi0: np.ndarray, steps: int = 20) -> LanczosResult:
    """`steps` Lanczos iterations from psi0 (host array); returns the lowest
    Ritz value and the normalised Ritz vector (on the device)."""
    n = hv.n
    vec = DeviceVectors(hv.dtype, n, hv.stream)
    esz = hv.esz
    basis = [capi.DeviceBuffer.from_numpy(psi0)]
    w = capi.DeviceBuffer(n * esz)
    lib = capi.lib()
    t0 = time.perf_counter()
    tmv = 0.0
    nrm = np.sqrt(vec.dot(basis[0].ptr, basis[0].ptr).real)
    vec.axpby(1.0 / nrm, basis[0].ptr, 0.0, basis[0].ptr)
    alphas, betas = [], []
    for j in range(steps):
        t1 = time.perf_counter()
        hv.apply(basis[j].ptr, w.ptr)
        capi.check(lib.qlt_stream_sync(hv.stream))
        tmv += time.perf_counter() - t1
        a = vec.dot(basis[j].ptr, w.ptr).real
        alphas.append(a)
        vec.axpby(-a, basis[j].ptr, 1.0, w.ptr)
        if j > 0:
            vec.axpby(-betas[-1], basis[j - 1].ptr, 1.0, w.ptr)
        for _ in range(2):                      # full reorthogonalisation
            for v in basis:
                c = vec.dot(v.ptr, w.ptr)
                vec.axpby(-c, v.ptr, 1.0, w.ptr)
        b = np.sqrt(vec.dot(w.ptr, w.ptr).real)
        if j == steps - 1 or b == 0.0:
            break
        betas.append(b)
        nxt = capi.DeviceBuffer(n * esz)
        capi.check(lib.qlt_memset_zero(nxt.ptr, n * esz, hv.stream))
        vec.axpby(1.0 / b, w.ptr, 0.0, nxt.ptr)
        basis.append(nxt)
    k = len(alphas)
    T = np.diag(alphas) + np.diag(betas[:k - 1], 1) + np.diag(betas[:k - 1], -1)
    evals, evecs = np.linalg.eigh(T)
    y = evecs[:, 0]
    gs = capi.DeviceBuffer(n * esz)
    capi.check(lib.qlt_memset_zero(gs.ptr, n * esz, hv.stream))
    for i in range(k):
        vec.axpby(float(y[i]), basis[i].ptr, 1.0, gs.ptr)
    nrm = np.sqrt(vec.dot(gs.ptr, gs.ptr).real)
    vec.axpby(1.0 / nrm, gs.ptr, 0.0, gs.ptr)
    capi.check(lib.qlt_stream_sync(hv.stream))
    return LanczosResult(float(evals[0]), np.array(alphas), np.array(betas), gs, time.perf_counter() - t0, tmv)


class KrylovArena:
    """Device buffers of a Lanczos run whose iteration launches with no host
    round trip: the basis arena (steps + 1 vectors), the active-vector count,
    the Gram-Schmidt coefficients of both passes, the Lanczos coefficients and
    the reduction workspace, all in device memory."""

    def __init__(self, hv: DeviceMatvec, steps: int):
        if hv.stream is None:
            raise ValueError("KrylovArena needs a matvec on a created stream (graph capture cannot use the null stream)")
        self.hv, self.steps = hv, steps
        self.k_max = steps + 1
        self.n, esz = hv.n, hv.esz
        self.dt = capi.QLT_C128 if hv.dtype == "c128" else capi.QLT_F64
        self.V = capi.DeviceBuffer(self.k_max * self.n * esz)
        self.x = capi.DeviceBuffer(self.n * esz)
        self.w = capi.DeviceBuffer(self.n * esz)
        self.count = capi.DeviceBuffer(4)
        self.c1 = capi.DeviceBuffer(16 * self.k_max)
        self.c2 = capi.DeviceBuffer(16 * self.k_max)
        self.alphas = capi.DeviceBuffer(8 * self.k_max)
        self.betas = capi.DeviceBuffer(8 * self.k_max)
        self.work = capi.DeviceBuffer(8 * (2 * 512 * self.k_max + 1))

    def reset(self, psi0: np.ndarray) -> None:
        """V_0 = psi0 / |psi0|, one active vector, coefficients cleared."""
        lib, st = capi.lib(), self.hv.stream
        v0 = np.ascontiguousarray(psi0 / np.sqrt(np.vdot(psi0, psi0).real))
        for buf in (self.V, self.alphas, self.betas, self.c1, self.c2):
            capi.check(lib.qlt_memset_zero(buf.ptr, buf.nbytes, st))
        capi.check(lib.qlt_stream_sync(st))
        capi.check(lib.qlt_memcpy_h2d(self.V.ptr, v0.ctypes.data_as(C.c_void_p), v0.nbytes))
        one = np.array([1], dtype=np.int32)
        capi.check(lib.qlt_memcpy_h2d(self.count.ptr, one.ctypes.data_as(C.c_void_p), 4))

    def iteration(self) -> None:
        """One Lanczos iteration, all on the matvec's stream: x = V_{m-1};
        w = H x; two passes of w -= V (V^H w) over the m active vectors (the
        first pass's <V_{m-1}, w> is alpha); beta = |w|; V_m = w / beta."""
        lib, st, n, k, dt = capi.lib(), self.hv.stream, self.n, self.k_max, self.dt
        capi.check(lib.qlt_basis_select(dt, n, k, self.count.ptr, -1, self.V.ptr, n, self.x.ptr, st))
        self.hv.apply(self.x.ptr, self.w.ptr)
        # pass 1: c1 = V^H w; w -= V c1 together with pass 2's c2 = V^H w
        # (one read of the basis for both); pass 2: w -= V c2
        capi.check(lib.qlt_basis_dot(dt, n, k, self.count.ptr, self.V.ptr, n, self.w.ptr, self.c1.ptr,
                                     self.work.ptr, st))
        capi.check(lib.qlt_basis_update_dot(dt, n, k, self.count.ptr, -1.0, self.c1.ptr, self.V.ptr, n, self.w.ptr,
                                            self.c2.ptr, self.work.ptr, st))
        capi.check(lib.qlt_basis_update(dt, n, k, self.count.ptr, -1.0, self.c2.ptr, self.V.ptr, n, self.w.ptr, st))
        capi.check(lib.qlt_lanczos_extend(dt, n, k, self.count.ptr, self.c1.ptr, self.w.ptr, self.V.ptr, n,
                                          self.alphas.ptr, self.betas.ptr, self.work.ptr, st))

    def coefficients(self):
        capi.check(capi.lib().qlt_stream_sync(self.hv.stream))
        return (self.alphas.to_numpy(np.float64)[:self.steps], self.betas.to_numpy(np.float64)[:self.steps])

    def ritz_vector(self, y: np.ndarray) -> capi.DeviceBuffer:
        """sum_i y_i V_i, normalised, on the device."""
        lib, st, n, dt = capi.lib(), self.hv.stream, self.n, self.dt
        k = len(y)
        gs = capi.DeviceBuffer(n * self.hv.esz)
        capi.check(lib.qlt_memset_zero(gs.ptr, gs.nbytes, st))
        coef = np.zeros(2 * k, dtype=np.float64)
        coef[0::2] = y
        cbuf = capi.DeviceBuffer.from_numpy(coef)
        cnt = capi.DeviceBuffer.from_numpy(np.array([k], dtype=np.int32))
        capi.check(lib.qlt_basis_update(dt, n, k, cnt.ptr, 1.0, cbuf.ptr, self.V.ptr, n, gs.ptr, st))
        vec = DeviceVectors(self.hv.dtype, n, st)
        nrm = np.sqrt(vec.dot(gs.ptr, gs.ptr).real)
        vec.axpby(1.0 / nrm, gs.ptr, 0.0, gs.ptr)
        capi.check(lib.qlt_stream_sync(st))
        return gs


def lanczos_graph(hv: DeviceMatvec, psi0: np.ndarray, steps: int = 20, use_graph: bool = True,
                  arena: "KrylovArena | None" = None) -> LanczosResult:
    """`steps` Lanczos iterations whose every vector operation stays on the
    device (KrylovArena).  With `use_graph` one iteration is captured once as
    a hipGraph and replayed `steps` times; otherwise the same launches run
    eagerly.  Classical Gram-Schmidt against the whole basis, twice, takes
    the place of the three-term recurrence plus reorthogonalisation (the same
    projections in exact arithmetic).  The tridiagonal matrix is read back
    once at the end and solved in FP64 on the host."""
    lib = capi.lib()
    ar = arena if arena is not None else KrylovArena(hv, steps)
    ar.reset(psi0)
    t0 = time.perf_counter()
    if use_graph:
        capi.check(lib.qlt_graph_begin(hv.stream))
        try:
            ar.iteration()
        finally:
            g = C.c_void_p()
            capi.check(lib.qlt_graph_end(hv.stream, C.byref(g)))
        try:
            for _ in range(steps):
                capi.check(lib.qlt_graph_replay(g, hv.stream))
            capi.check(lib.qlt_stream_sync(hv.stream))
        finally:
            capi.check(lib.qlt_graph_destroy(g))
    else:
        for _ in range(steps):
            ar.iteration()
        capi.check(lib.qlt_stream_sync(hv.stream))
    iter_s = time.perf_counter() - t0
    alphas, betas = ar.coefficients()
    k = steps
    for j in range(steps - 1):        # an invariant subspace ends the recurrence
        if betas[j] == 0.0:
            k = j + 1
            break
    T = np.diag(alphas[:k]) + np.diag(betas[:k - 1], 1) + np.diag(betas[:k - 1], -1)
    evals, evecs = np.linalg.eigh(T)
    gs = ar.ritz_vector(evecs[:, 0])
    return LanczosResult(float(evals[0]), alphas[:k], betas[:k - 1], gs, time.perf_counter() - t0, iter_s)
