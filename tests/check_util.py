"""Planted-invalid match arrays for the result-check tests (test helper)."""
import numpy as np


def plant_invalid(m, csr, rng):
    """A copy of match array m with every kind of defect testMatching counts
    (src/matchcpu.cu:127-246): matched vertices whose partner is reset,
    a pair of non-neighbours, a third member of a pair, and values outside the
    id range: one alone, two neighbours sharing one (testMatching does not
    read the value itself, so that pair is valid to it) and three sharing one."""
    m = np.asarray(m, np.int32).copy()
    n = csr.n
    matched = np.flatnonzero(m >= 4)
    free = np.flatnonzero(m < 4)
    if matched.size >= 8:
        m[rng.choice(matched, max(1, matched.size // 50), replace=False)] = 1
    if free.size >= 6:
        far = rng.choice(free, 6, replace=False)
        m[far[:2]] = 4 + int(far[:2].min())          # a pair that are (most likely) not neighbours
        m[far[2]] = 4 + n + 9                         # alone, outside the range
        m[far[3:6]] = 4 + n + 11                      # three outside the range
    matched = np.flatnonzero(m >= 4)
    if matched.size:
        third = rng.choice(np.flatnonzero(m < 4)) if np.any(m < 4) else None
        if third is not None:
            m[third] = m[matched[0]]                  # a third member
    # two neighbours sharing one value outside the range
    for v in rng.permutation(n)[: min(n, 64)]:
        a, b = csr.ranges[v]
        if b > a and csr.neighbours[a] != v:
            m[v] = m[csr.neighbours[a]] = 4 + n + 13
            break
    return m
