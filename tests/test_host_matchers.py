"""The reference's sequential host matchers (types 0-5, src/matchcpu.cu:259-810)
as the library runs them (mm_host_match) against the oracle's restatements,
and the oracle's types 1-3 pinned by closed forms: maximum matchings on
forests for Karp-Sipser, paths / cycles / stars / complete bipartite graphs
and a hand-worked case for the minimum degree matchers.  Host only."""
import numpy as np
import pytest

import mmamd
import oracle
from mmamd import CSR


def csr_of(n, edges, weights=None):
    return CSR.from_edges(n, np.asarray(edges, np.int64).reshape(-1, 2), weights)


def random_graph(n, m, seed, isolated=0):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, n - isolated, 3 * m)
    b = rng.integers(0, n - isolated, 3 * m)
    keep = a != b
    e = np.unique(np.sort(np.stack([a[keep], b[keep]], 1), axis=1), axis=0)
    e = e[rng.permutation(len(e))[:m]]
    w = rng.uniform(0.01, 1.0, len(e)).astype(np.float32)
    return csr_of(n, e, w)


def random_forest(n, seed, p_root=0.1):
    """Each vertex v > 0 hangs off a random earlier vertex, or starts a new tree."""
    rng = np.random.default_rng(seed)
    parent = [-1] + [int(rng.integers(0, v)) if rng.uniform() > p_root else -1 for v in range(1, n)]
    edges = [(p, v) for v, p in enumerate(parent) if p >= 0]
    perm = rng.permutation(n)  # shuffled ids, so leaves are not in any particular order
    return csr_of(n, [(perm[a], perm[b]) for a, b in edges] or np.zeros((0, 2))), parent


def forest_maximum_matching(parent):
    """Maximum matching size of a forest given by parent pointers (children
    after parents): tree DP with states (best with v free, best overall)."""
    n = len(parent)
    free = [0] * n  # best in v's subtree with v unmatched
    best = [0] * n  # best in v's subtree
    for v in range(n - 1, -1, -1):
        kids = [c for c in range(v + 1, n) if parent[c] == v]
        f = sum(best[c] for c in kids)
        b = f
        for c in kids:
            b = max(b, f - best[c] + free[c] + 1)
        free[v], best[v] = f, b
    return sum(best[v] for v in range(n) if parent[v] < 0)


def pairs(match):
    m = np.asarray(match)
    return int((m >= 4).sum()) // 2


def assert_valid_maximal(csr, m):
    ok, under, multi, invalid, matchable = oracle.test_matching(csr, m)
    assert ok and under == multi == invalid == 0
    assert matchable == 0


GRAPHS = [
    ("random500", lambda: random_graph(500, 1200, 1)),
    ("random2000_isolated", lambda: random_graph(2000, 3000, 2, isolated=50)),
    ("sparse3000", lambda: random_graph(3000, 1600, 3)),
    ("grid", lambda: mmamd.gen_grid(40, 30, seed=4, weighted=True)),
    ("rgg", lambda: mmamd.gen_rgg(3000, seed=5, weighted=True)),
    ("rmat", lambda: mmamd.gen_rmat(11, seed=6, weighted=True)),
    ("forest", lambda: random_forest(2500, 7)[0]),
]

HOST_ORACLE = {
    mmamd.HOST_RANDOM: oracle.cpu_random,
    mmamd.HOST_STAT_MIN_DEG: oracle.cpu_stat_min_deg,
    mmamd.HOST_KARP_SIPSER: oracle.cpu_karp_sipser,
    mmamd.HOST_MIN_DEG: oracle.cpu_min_deg,
    mmamd.HOST_WEIGHTED: oracle.cpu_weighted,
    mmamd.HOST_WEIGHTED_EDGE: oracle.cpu_weighted_edge,
}


@pytest.mark.parametrize("name,make", GRAPHS, ids=[g[0] for g in GRAPHS])
def test_library_host_matchers_equal_the_oracle(name, make):
    csr = make()
    for t, fn in HOST_ORACLE.items():
        got = mmamd.host_match(csr, t)
        np.testing.assert_array_equal(got, fn(csr), err_msg=f"type {t} on {name}")
        assert_valid_maximal(csr, got)


def test_karp_sipser_is_maximum_on_forests():
    for seed in range(12):
        csr, parent = random_forest(400 + 37 * seed, 100 + seed, p_root=0.02 + 0.03 * (seed % 4))
        m = oracle.cpu_karp_sipser(csr)
        assert_valid_maximal(csr, m)
        assert pairs(m) == forest_maximum_matching(parent), seed


def test_karp_sipser_known_answer():
    # 0-1-2-3 path plus triangle 4-5-6 hanging off 3 by the edge 3-4.
    # Degrees 1,2,2,2,3,2,2: singleton 0 takes 1, lowering 2 to 1; 2 takes 3,
    # lowering 4 to 2.  The pass then reaches 4 (deg 2): first free neighbour 5.
    csr = csr_of(7, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (5, 6), (4, 6)])
    assert oracle.cpu_karp_sipser(csr).tolist() == [4, 4, 6, 6, 8, 8, 0]


@pytest.mark.parametrize("fn", [oracle.cpu_min_deg, oracle.cpu_stat_min_deg, oracle.cpu_karp_sipser])
def test_degree_matchers_closed_forms(fn):
    for n in (2, 3, 7, 10, 31):
        path = csr_of(n, [(i, i + 1) for i in range(n - 1)])
        m = fn(path)
        assert_valid_maximal(path, m)
        if fn is not oracle.cpu_stat_min_deg:  # the static order can split a path badly
            assert pairs(m) == n // 2, n
    for n in (4, 9, 12):
        cyc = csr_of(n, [(i, (i + 1) % n) for i in range(n)])
        m = fn(cyc)
        assert_valid_maximal(cyc, m)
    star = csr_of(9, [(0, k) for k in range(1, 9)])
    assert pairs(fn(star)) == 1
    for a, b in ((3, 5), (4, 4), (6, 2)):
        kab = csr_of(a + b, [(i, a + j) for i in range(a) for j in range(b)])
        m = fn(kab)
        assert_valid_maximal(kab, m)
        assert pairs(m) == min(a, b)


def test_min_deg_maximum_on_paths_and_cycles():
    for n in (5, 8, 13, 40):
        cyc = csr_of(n, [(i, (i + 1) % n) for i in range(n)])
        assert pairs(oracle.cpu_min_deg(cyc)) == n // 2


def test_min_deg_known_answer_uses_recency_among_equal_degrees():
    # Square 0-1-2-3-0 with a pendant 4 on 0.  Degrees 3,2,2,2,1: 4 goes first
    # and takes 0; 1 and 3 drop to degree 1 and are re-inserted in row order
    # of 0 (1 then 3), ahead of 2 (degree 2).  So 1 is next and takes 2, which
    # drops 3 to degree 0: it is left alone.
    csr = csr_of(5, [(0, 1), (1, 2), (2, 3), (3, 0), (0, 4)])
    assert oracle.cpu_min_deg(csr).tolist() == [4, 5, 5, 0, 4]
    # Same square, pendant on 1 instead: 4 takes 1; 0 and 2 drop to 1 and
    # re-enter in row order of 1 (0 then 2): 0 takes 3, leaving 2.
    csr = csr_of(5, [(0, 1), (1, 2), (2, 3), (3, 0), (1, 4)])
    assert oracle.cpu_min_deg(csr).tolist() == [4, 5, 0, 4, 5]


def test_stat_min_deg_known_answer():
    # Path 0-1-2-3: vertices 0 and 3 have degree 1 and come first in either
    # order; both take their only neighbour, a perfect matching.
    csr = csr_of(4, [(0, 1), (1, 2), (2, 3)])
    assert oracle.cpu_stat_min_deg(csr).tolist() == [4, 4, 6, 6]
    # Star with centre 0 and a tail 1-5: the leaves 2, 3, 4 and 5 (degree 1)
    # come first; the first of them takes the centre, and the tail edge 1-5
    # is matched from 5.
    csr = csr_of(6, [(0, 1), (0, 2), (0, 3), (0, 4), (1, 5)])
    m = oracle.cpu_stat_min_deg(csr)
    assert pairs(m) == 2 and m[1] == m[5] == 5


def test_host_match_rejects_bad_input():
    csr = random_graph(50, 60, 9)
    with pytest.raises(mmamd.MatchError):
        mmamd.host_match(csr, 6)
    bad = CSR(csr.ranges.copy(), csr.neighbours.copy() + 1000, csr.weights)
    with pytest.raises(mmamd.MatchError):
        mmamd.host_match(bad, 0)


def test_library_result_check_equals_the_oracle_on_planted_invalid_arrays():
    """mm_test_matching / mm_get_weight (the drop-in's GraphMatching::testMatching
    and getWeight) against the oracle, on valid results and on arrays with every
    defect testMatching counts, values outside the id range included."""
    from check_util import plant_invalid

    rng = np.random.default_rng(23)
    seeds = mmamd.reference_seeds(255)
    for csr in (mmamd.gen_rgg(3000, seed=41, weighted=True), mmamd.gen_grid(40, 30, seed=42, weighted=True),
                mmamd.gen_rmat(11, seed=43, weighted=True)):
        m = oracle.match(csr, True, True, seeds)[0]
        for arr in (m, plant_invalid(m, csr, rng), plant_invalid(m, csr, rng)):
            assert mmamd.test_matching(csr, arr) == oracle.test_matching(csr, arr)
            assert mmamd.get_weight(csr, arr) == oracle.get_weight(csr, arr)
        bad = plant_invalid(m, csr, rng)
        exp = oracle.test_matching(csr, bad)
        assert not exp[0] and exp[1] and exp[2] and exp[3]
