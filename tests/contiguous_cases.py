"""Contiguous-axes contraction cases (ContractContiguousAxes,
contract_contiguous_axes.h:305-353): contracted axes (start + i) % rank, saved
axes cyclic from the end of the run, (end + i) % rank."""
from __future__ import annotations

import numpy as np

from tensortoolkit_amd.blocksparse import BlockSparseTensor
from tests.helpers import random_index


def contiguous_axes(rank, start, size):
    ctr = [(start + i) % rank for i in range(size)]
    end = (start + size) % rank
    saved = [(end + i) % rank for i in range(rank - size)]
    return ctr, saved


# (rank A, rank B, A start, B start, size): runs at the tail, head, middle and
# wrapping round the end, so that the cyclic saved order differs from the
# ascending one in most cases
SHAPES = [
    (3, 3, 2, 0, 1), (3, 3, 1, 0, 1), (3, 3, 0, 2, 1), (3, 3, 1, 1, 1),
    (4, 3, 1, 0, 2), (4, 4, 3, 2, 2), (4, 4, 2, 3, 2), (4, 3, 3, 2, 2),
    (3, 4, 2, 3, 2), (5, 4, 3, 1, 3), (4, 4, 1, 1, 3), (3, 3, 1, 2, 3),
    (4, 4, 1, 2, 2), (5, 4, 2, 1, 2), (4, 4, 2, 1, 1), (5, 3, 1, 1, 2),
]


def contiguous_case(rng, shape, dtype="f64", nq=1, max_deg=6, min_deg=1, tries=400):
    ra, rb, sa, sb, size = shape
    a_ctr, a_sv = contiguous_axes(ra, sa, size)
    b_ctr, b_sv = contiguous_axes(rb, sb, size)
    for _ in range(tries):
        a_idx = [random_index(rng, nq, 3, max_deg, min_deg=min_deg) for _ in range(ra)]
        b_idx = [None] * rb
        for x, y in zip(a_ctr, b_ctr):
            b_idx[y] = a_idx[x].inverse()
        for j in range(rb):
            if b_idx[j] is None:
                b_idx[j] = random_index(rng, nq, 3, max_deg, min_deg=min_deg)
        div_a = tuple(int(v) for v in rng.integers(-1, 2, size=nq))
        div_b = tuple(int(v) for v in rng.integers(-1, 2, size=nq))
        a = BlockSparseTensor.allowed_blocks(a_idx, div_a, dtype)
        b = BlockSparseTensor.allowed_blocks(b_idx, div_b, dtype)
        if a.n_blks == 0 or b.n_blks == 0:
            continue
        if not {tuple(c[a_ctr]) for c in a.coords} & {tuple(c[b_ctr]) for c in b.coords}:
            continue
        seed = int(rng.integers(1, 2 ** 62))
        a.fill_uniform(seed)
        b.fill_uniform(seed + 29)
        return a, b, (a_ctr, b_ctr, a_sv, b_sv)
    raise RuntimeError("could not draw a non-empty contiguous case")


def dense_reference(a, b, axes):
    """np.tensordot of the densified tensors, its free axes put in the saved
    orders: tensordot leaves A's free axes ascending, then B's."""
    a_ctr, b_ctr, a_sv, b_sv = axes
    d = np.tensordot(a.to_dense(), b.to_dense(), axes=(a_ctr, b_ctr))
    a_free = sorted(a_sv)
    b_free = sorted(b_sv)
    perm = [a_free.index(x) for x in a_sv] + [len(a_free) + b_free.index(x) for x in b_sv]
    return np.transpose(d, perm) if perm else d
