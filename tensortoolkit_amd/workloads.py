"""Synthetic workloads of the block-sparse contraction benchmark.

Index conventions (fixed here, used by the tests and the benchmark):

* rank-3 U(1) pair: A, B = (l: IN, p: IN, r: OUT), div 0, contract A.r with B.l.
* two-site effective-Hamiltonian matvec, psi' = L . psi . W1 . W2 . R:
    psi (l: IN, s1: IN, s2: IN, r: OUT)          div 0
    L   (a': IN, w: OUT, a: OUT)                 div 0, a contracts psi.l
    W   (wl: IN, wr: OUT, so: IN, si: OUT)       div 0, si contracts a physical leg
    R   (b: IN, w: IN, b': OUT)                  div 0, b contracts psi.r
  chain (axes as passed to Contract):
    T1 = Contract(L,  psi, {{2},{0}})      -> (a', w, s1, s2, r)
    T2 = Contract(T1, W1,  {{1,2},{0,3}})  -> (a', s2, r, wr, so1)
    T3 = Contract(T2, W2,  {{1,3},{3,0}})  -> (a', r, so1, wr2, so2)
    T4 = Contract(T3, R,   {{1,3},{0,1}})  -> (a', so1, so2, b')  (psi's structure)
  The MPO bond carries operator charges: Heisenberg {-1:1, 0:3, +1:1};
  U(1)xU(1) Hubbard {(0,0):2, (1,1):1, (1,-1):1, (-1,-1):1, (-1,1):1}.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .blocksparse import IN, OUT, BlockSparseTensor, Index

MATVEC_AXES = [
    ((2,), (0,)),
    ((1, 2), (0, 3)),
    ((1, 3), (3, 0)),
    ((1, 3), (0, 1)),
]


def largest_remainder(weights, total: int) -> np.ndarray:
    """Integers proportional to `weights` summing exactly to `total` (each >= 1)."""
    w = np.asarray(weights, dtype=np.float64)
    if total < len(w):
        raise ValueError(f"cannot give each of {len(w)} sectors a degeneracy >= 1 from a total of {total}")
    raw = w / w.sum() * total
    base = np.floor(raw).astype(np.int64)
    base = np.maximum(base, 1)
    rem = total - base.sum()
    frac = raw - np.floor(raw)
    order = np.argsort(-frac, kind="stable")
    i = 0
    while rem > 0:
        base[order[i % len(order)]] += 1
        rem -= 1
        i += 1
    while rem < 0:
        j = order[::-1][i % len(order)]
        if base[j] > 1:
            base[j] -= 1
            rem += 1
        i += 1
    return base


def u1_gaussian_degs(qmax: int, total: int, var: float = 9.0):
    charges = np.arange(-qmax, qmax + 1)
    return charges, largest_remainder(np.exp(-charges.astype(float) ** 2 / (2 * var)), total)


def u1u1_gaussian_degs(nmax: int, total: int, var: float = 9.0):
    qns = [(n, s) for n in range(-nmax, nmax + 1) for s in range(-nmax, nmax + 1) if (n - s) % 2 == 0]
    w = [np.exp(-(n * n + s * s) / (2 * var)) for n, s in qns]
    return qns, largest_remainder(w, total)


@dataclass
class Matvec:
    """Tensors of one two-site matvec and their element type."""

    L: BlockSparseTensor
    psi: BlockSparseTensor
    W: BlockSparseTensor
    R: BlockSparseTensor
    dtype: str
    name: str


def _vindex(qns, degs, direction) -> Index:
    return Index(tuple(tuple(q) if isinstance(q, tuple) else (int(q),) for q in qns),
                 tuple(int(d) for d in degs), direction)


def build_matvec(virt_qns, virt_degs, phys_qns, mpo_qns, mpo_degs, dtype: str, seed: int,
                 name: str) -> Matvec:
    zero = tuple([0] * len(virt_qns[0] if isinstance(virt_qns[0], tuple) else (0,)))
    v_in = _vindex(virt_qns, virt_degs, IN)
    v_out = _vindex(virt_qns, virt_degs, OUT)
    p_in = _vindex(phys_qns, [1] * len(phys_qns), IN)
    p_out = _vindex(phys_qns, [1] * len(phys_qns), OUT)
    w_in = _vindex(mpo_qns, mpo_degs, IN)
    w_out = _vindex(mpo_qns, mpo_degs, OUT)
    psi = BlockSparseTensor.allowed_blocks([v_in, p_in, p_in, v_out], zero, dtype).fill_uniform(seed)
    L = BlockSparseTensor.allowed_blocks([v_in, w_out, v_out], zero, dtype).fill_uniform(seed + 1)
    W = BlockSparseTensor.allowed_blocks([w_in, w_out, p_in, p_out], zero, dtype).fill_uniform(seed + 2)
    R = BlockSparseTensor.allowed_blocks([v_in, w_in, v_out], zero, dtype).fill_uniform(seed + 3)
    return Matvec(L=L, psi=psi, W=W, R=R, dtype=dtype, name=name)


def config0(seed: int = 20261202, dtype: str = "f64"):
    """Rank-3 U(1) pair: charges -4..4 x deg 16 on virtual legs, {0,1} x 1 physical."""
    charges = list(range(-4, 5))
    l_in = Index(tuple((q,) for q in charges), tuple([16] * 9), IN)
    r_out = Index(tuple((q,) for q in charges), tuple([16] * 9), OUT)
    p_in = Index(((0,), (1,)), (1, 1), IN)
    a = BlockSparseTensor.allowed_blocks([l_in, p_in, r_out], (0,), dtype).fill_uniform(seed)
    b = BlockSparseTensor.allowed_blocks([l_in, p_in, r_out], (0,), dtype).fill_uniform(seed + 1)
    return a, b, ((2,), (0,))


HEIS_MPO_QNS = [(-1,), (0,), (1,)]
HEIS_MPO_DEGS = [1, 3, 1]
HUBB_MPO_QNS = [(0, 0), (1, 1), (1, -1), (-1, -1), (-1, 1)]
HUBB_MPO_DEGS = [2, 1, 1, 1, 1]


def config1(seed: int = 20261202, D: int = 2048, dtype: str = "f64") -> Matvec:
    """D=2048 U(1) two-site matvec: charges -8..8, deg(q) ~ exp(-q^2/18);
    L and R symmetrised to be Hermitian (L[a',w,a] = conj L[a,w~,a'],
    likewise R), as environments of a Hermitian Hamiltonian are."""
    charges, degs = u1_gaussian_degs(8, D)
    mv = build_matvec([(int(q),) for q in charges], degs, [(0,), (1,)], HEIS_MPO_QNS, HEIS_MPO_DEGS,
                      dtype, seed, f"u1_matvec_D{D}")
    wbar_l = _partner_sector(mv.L.indexes[1])
    wbar_r = _partner_sector(mv.R.indexes[1])
    mv.L = _symmetrize(mv.L, lambda c: (c[2], int(wbar_l[c[1]]), c[0]), (2, 1, 0))
    mv.R = _symmetrize(mv.R, lambda c: (c[2], int(wbar_r[c[1]]), c[0]), (2, 1, 0))
    return mv


def config2(seed: int = 20261202, dtype: str = "f64") -> Matvec:
    """Lognormal(mu=4, sigma=1) sector degeneracies, clipped to 1..512, charges -8..8."""
    rng = np.random.default_rng(seed)
    charges = np.arange(-8, 9)
    degs = np.clip(np.rint(rng.lognormal(4.0, 1.0, size=len(charges))), 1, 512).astype(np.int64)
    return build_matvec([(int(q),) for q in charges], degs, [(0,), (1,)], HEIS_MPO_QNS, HEIS_MPO_DEGS,
                        dtype, seed, f"u1_lognormal_D{int(degs.sum())}")


def config3(seed: int = 20261202, D: int = 4096, dtype: str = "c128") -> Matvec:
    """U(1)xU(1) (N, 2Sz) Hubbard-like matvec: 85 sectors N,2Sz in -6..6 (equal parity)."""
    qns, degs = u1u1_gaussian_degs(6, D)
    phys = [(0, 0), (1, 1), (1, -1), (2, 0)]
    return build_matvec(qns, degs, phys, HUBB_MPO_QNS, HUBB_MPO_DEGS, dtype, seed, f"u1u1_matvec_D{D}")


def small_matvec(seed: int = 7, D: int = 96, dtype: str = "f64", u1u1: bool = False) -> Matvec:
    """Scaled-down matvec with the same structure (parity tests)."""
    if u1u1:
        qns, degs = u1u1_gaussian_degs(2, D, var=2.0)
        return build_matvec(qns, degs, [(0, 0), (1, 1), (1, -1), (2, 0)], HUBB_MPO_QNS, HUBB_MPO_DEGS,
                            dtype, seed, f"small_u1u1_D{D}")
    charges, degs = u1_gaussian_degs(3, D, var=3.0)
    return build_matvec([(int(q),) for q in charges], degs, [(0,), (1,)], HEIS_MPO_QNS, HEIS_MPO_DEGS,
                        dtype, seed, f"small_u1_D{D}")


def _partner_sector(ix: Index) -> np.ndarray:
    """Sector of -q for every sector q of an MPO bond index (the conjugate
    operator's bond sector, e.g. S+ <-> S-)."""
    pos = {tuple(q): k for k, q in enumerate(ix.qns)}
    out = np.empty(ix.nsec, dtype=np.int64)
    for k, q in enumerate(ix.qns):
        partner = pos[tuple(-v for v in q)]
        assert ix.degs[partner] == ix.degs[k], "conjugate MPO sectors need equal degeneracies"
        out[k] = partner
    return out


def _symmetrize(t: BlockSparseTensor, coord_map, axes_perm) -> BlockSparseTensor:
    """t'[block c] = (t[c] + conj(t[pi(c)]).transpose(axes_perm)) / 2, with
    pi(c) = coord_map(c); the partner block must exist."""
    lut = {tuple(int(x) for x in t.coords[i]): i for i in range(t.n_blks)}
    data = np.empty_like(t.data)
    for i in range(t.n_blks):
        j = lut[coord_map(tuple(int(x) for x in t.coords[i]))]
        x = t.block(i)
        y = np.conj(t.block(j)).transpose(axes_perm)
        data[t.offsets[i]:t.offsets[i] + x.size] = (0.5 * (x + y)).reshape(-1)
    return BlockSparseTensor(t.indexes, t.coords, t.dims, dtype=t.dtype, data=data, offsets=t.offsets)


def hermitian_matvec(mv: Matvec) -> Matvec:
    """The same tensors made to give a Hermitian effective Hamiltonian
    H = sum L_w (x) W_ww' (x) W_w'w'' (x) R_w'': with w~ the conjugate MPO
    sector, L[a',w,a] = conj L[a,w~,a'], R[b,w,b'] = conj R[b',w~,b] and
    W[wl,wr,so,si] = conj W[wl~,wr~,si,so], each imposed by averaging a block
    with its conjugate partner.  Every term of H is then a product of
    operators whose sum over the bond indices is Hermitian."""
    wbar_l = _partner_sector(mv.L.indexes[1])
    wbar_r = _partner_sector(mv.R.indexes[1])
    wbar_w = _partner_sector(mv.W.indexes[0])
    L = _symmetrize(mv.L, lambda c: (c[2], int(wbar_l[c[1]]), c[0]), (2, 1, 0))
    R = _symmetrize(mv.R, lambda c: (c[2], int(wbar_r[c[1]]), c[0]), (2, 1, 0))
    W = _symmetrize(mv.W, lambda c: (int(wbar_w[c[0]]), int(wbar_w[c[1]]), c[3], c[2]), (0, 1, 3, 2))
    return Matvec(L=L, psi=mv.psi, W=W, R=R, dtype=mv.dtype, name=mv.name + "_hermitian")


def config4(seed: int = 20261202, D: int = 4096) -> Matvec:
    """End-to-end two-site update on the U(1)xU(1) D=4096 complex128 tensors
    with L, W, R made Hermitian (Lanczos needs a Hermitian H_eff)."""
    return hermitian_matvec(config3(seed, D))
