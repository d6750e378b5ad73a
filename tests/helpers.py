"""Random block-sparse contraction cases for the parity tests."""
from __future__ import annotations

import numpy as np

from tensortoolkit_amd.blocksparse import IN, OUT, BlockSparseTensor, Index


def random_index(rng, nq=1, max_sectors=4, max_deg=5, direction=None, min_deg=1) -> Index:
    nsec = int(rng.integers(1, max_sectors + 1))
    if nq == 1:
        qs = sorted(rng.choice(np.arange(-3, 4), size=nsec, replace=False).tolist())
        qns = tuple((int(q),) for q in qs)
    else:
        pool = [(a, b) for a in range(-2, 3) for b in range(-2, 3)]
        sel = rng.choice(len(pool), size=min(nsec, len(pool)), replace=False)
        qns = tuple(pool[i] for i in sorted(sel))
    degs = tuple(int(d) for d in rng.integers(min_deg, max_deg + 1, size=len(qns)))
    if direction is None:
        direction = IN if rng.random() < 0.5 else OUT
    return Index(qns, degs, direction)


def random_case(rng, dtype="f64", nq=1, max_rank=4, max_sectors=4, max_deg=5, allow_scalar=True,
                min_blocks=1, tries=400, min_deg=1, need_tasks=True):
    """(A, B, a_axes, b_axes) with A[a_i] == InverseIndex(B[b_i]); random ranks,
    contracted subsets in random order, random divergences."""
    for _ in range(tries):
        ra = int(rng.integers(1, max_rank + 1))
        rb = int(rng.integers(1, max_rank + 1))
        n_ctr = int(rng.integers(1, min(ra, rb) + 1))
        if not allow_scalar and n_ctr == ra and n_ctr == rb:
            continue
        a_idx = [random_index(rng, nq, max_sectors, max_deg, min_deg=min_deg) for _ in range(ra)]
        a_ax = [int(x) for x in rng.permutation(ra)[:n_ctr]]
        b_ax = [int(x) for x in rng.permutation(rb)[:n_ctr]]
        b_idx = [None] * rb
        for x, y in zip(a_ax, b_ax):
            b_idx[y] = a_idx[x].inverse()
        for j in range(rb):
            if b_idx[j] is None:
                b_idx[j] = random_index(rng, nq, max_sectors, max_deg, min_deg=min_deg)
        div_a = tuple(int(v) for v in rng.integers(-1, 2, size=nq))
        div_b = tuple(int(v) for v in rng.integers(-1, 2, size=nq))
        a = BlockSparseTensor.allowed_blocks(a_idx, div_a, dtype)
        b = BlockSparseTensor.allowed_blocks(b_idx, div_b, dtype)
        if a.n_blks < min_blocks or b.n_blks < min_blocks:
            continue
        if need_tasks:
            ka = {tuple(c[a_ax]) for c in a.coords}
            kb = {tuple(c[b_ax]) for c in b.coords}
            if not ka & kb:
                continue
        seed = int(rng.integers(1, 2 ** 62))
        a.fill_uniform(seed)
        b.fill_uniform(seed + 17)
        return a, b, a_ax, b_ax
    raise RuntimeError("could not draw a non-empty case")


def rel_frobenius(x, ref) -> float:
    x = np.asarray(x)
    ref = np.asarray(ref)
    den = np.linalg.norm(ref.ravel())
    num = np.linalg.norm((x - ref).ravel())
    return float(num / den) if den > 0 else float(num)


def per_block_rel_err(c_tensor, c_ref_data) -> float:
    """Maximum over output blocks of the relative Frobenius error."""
    worst = 0.0
    for i in range(c_tensor.n_blks):
        size = int(np.prod(c_tensor.dims[i]))
        o = c_tensor.offsets[i]
        worst = max(worst, rel_frobenius(c_tensor.data[o:o + size], c_ref_data[o:o + size]))
    return worst
