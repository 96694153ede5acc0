"""Shared helpers of the tests: layout conversions between deal.II's full node set and
the stored (periodic images omitted) layout, and seeded random fields."""

from __future__ import annotations

import numpy as np


def full_shape(info):
    return tuple(reversed([info.degree * info.cells[d] + 1 for d in range(info.dim)]))


def stored_shape(info):
    return tuple(reversed([info.nodes[d] for d in range(info.dim)]))


def to_stored(full: np.ndarray, info) -> np.ndarray:
    """Drops periodic image planes: full lattice -> stored layout (flat)."""
    a = np.asarray(full).reshape(full_shape(info))
    sl = tuple(slice(0, info.nodes[info.dim - 1 - ax]) for ax in range(info.dim))
    return np.ascontiguousarray(a[sl]).ravel()


def to_full(stored: np.ndarray, info) -> np.ndarray:
    """Stored layout -> full lattice, images copied from their masters (flat)."""
    a = np.asarray(stored).reshape(stored_shape(info))
    for ax in range(info.dim):
        d = info.dim - 1 - ax
        if info.periodic[d]:
            first = np.take(a, [0], axis=ax)
            a = np.concatenate([a, first], axis=ax)
    return np.ascontiguousarray(a).ravel()


def random_stored(info, rng, lo=-1.0, hi=1.0, n_block=1):
    return rng.uniform(lo, hi, size=n_block * info.n_nodes)


def block_to_full(stored_block, info, n_block):
    n = info.n_nodes
    return np.concatenate([to_full(stored_block[k * n:(k + 1) * n], info) for k in range(n_block)])


def block_to_stored(full_block, info, n_block):
    nf = info.n_dofs
    return np.concatenate([to_stored(full_block[k * nf:(k + 1) * nf], info) for k in range(n_block)])


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
