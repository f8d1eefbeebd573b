"""Seeded synthetic workloads of the matcher's benchmarks (host side).

  rgg(log2n)     random geometric graph, n = 2^log2n points in the unit square,
                 r = 0.55 sqrt(ln n / n), weights (k + 1) / 2^24 per edge
  rmat(scale)    R-MAT, edge factor 16, (a, b, c, d) = (0.57, 0.19, 0.19, 0.05),
                 symmetrised, deduplicated, self-loops removed, same weights
  batch(count)   `count` independent graphs with n log-uniform in [2^8, 2^16]:
                 even k a near-square 2-D grid with the 4-neighbour stencil,
                 odd k an RGG as above; every edge weighted

Graphs are generated by the library's C generators (mm_gen_*); this module
only chooses their sizes and seeds and concatenates batches.
"""
from __future__ import annotations

import math

import numpy as np

import mmamd


def rgg(log2n: int = 22, seed: int = 12345) -> mmamd.CSR:
    return mmamd.gen_rgg(1 << log2n, seed=seed, weighted=True)


def rmat(scale: int = 24, seed: int = 12345) -> mmamd.CSR:
    return mmamd.gen_rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=seed, weighted=True)


def batch_sizes(count: int = 4096, min_log2n: int = 8, max_log2n: int = 16, seed: int = 12345):
    """(kind, n, graph seed) per graph: kind 'grid' (even k) or 'rgg' (odd k)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(round(2.0 ** rng.uniform(min_log2n, max_log2n)))
        gseed = int(rng.integers(1, 2**62))
        out.append(("grid" if k % 2 == 0 else "rgg", n, gseed))
    return out


def batch_graph(kind: str, n: int, gseed: int) -> mmamd.CSR:
    if kind == "grid":
        rows = max(1, int(math.isqrt(n)))
        cols = max(1, n // rows)
        return mmamd.gen_grid(rows, cols, seed=gseed, weighted=True)
    return mmamd.gen_rgg(n, seed=gseed, weighted=True)


def batch(count: int = 4096, min_log2n: int = 8, max_log2n: int = 16, seed: int = 12345) -> list:
    return [batch_graph(*spec) for spec in batch_sizes(count, min_log2n, max_log2n, seed)]


def concat(graphs: list) -> tuple[mmamd.CSR, np.ndarray]:
    """One CSR with global ids holding every graph (for host-side validity
    checks of a batch result) and the vertex offsets."""
    vo = np.zeros(len(graphs) + 1, np.int64)
    eo = np.zeros(len(graphs) + 1, np.int64)
    for k, g in enumerate(graphs):
        vo[k + 1] = vo[k] + g.n
        eo[k + 1] = eo[k] + g.nnz
    ranges = np.concatenate([g.ranges.astype(np.int64) + eo[k] for k, g in enumerate(graphs)])
    nb = np.concatenate([g.neighbours.astype(np.int64) + vo[k] for k, g in enumerate(graphs)])
    w = np.concatenate([g.weights for g in graphs])
    return mmamd.CSR(ranges.astype(np.int32), nb.astype(np.int32), w), vo


def batch_match_global(match_local: np.ndarray, vertex_offsets: np.ndarray) -> np.ndarray:
    """Batch results hold graph-local match values (4 + local id); shifts them
    to the concatenated CSR's global ids."""
    m = np.asarray(match_local, np.int64).copy()
    owner = np.repeat(np.arange(len(vertex_offsets) - 1), np.diff(vertex_offsets))
    matched = m >= 4
    m[matched] += vertex_offsets[owner[matched]]
    return m.astype(np.int32)
