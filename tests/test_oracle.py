"""Pins the CPU oracle (oracle/mm_oracle.c) before it is trusted as the checker.

Each check is chosen so a plausible mistake in the restatement (a wrong
constant, rotation, index, sentinel or tie rule) fails at least one of them:
  * hash: constants re-derived from mathematics, an independent numpy
    restatement over ~1e6 pairs, committed vectors, special cases;
  * rounds: committed tiny fixtures from the literal Python statement, closed
    forms (single edge, red-centred star, ties), the unit-weight identity
    weighted == random, invariants (valid, maximal at convergence), and
    serial == OpenMP;
  * metrics and CPU baselines: hand-computed known answers and the 1/2 bound
    against brute-force optimal matchings;
  * seeds: glibc rand() against the published TYPE_3 additive-feedback
    algorithm.
"""
import ctypes
import itertools
import json
import math
import os
import random

import numpy as np
import pytest

import literal
import oracle
from mmamd import CSR, reference_seeds

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def load_cases():
    with open(os.path.join(GOLD, "tiny_cases.json")) as f:
        return json.load(f)


def case_csr(c):
    w = np.array(c["weight_bits"], dtype=np.uint32).view(np.float32) if c["weight_bits"] else np.zeros(0, np.float32)
    return CSR.from_offsets(c["offsets"], np.array(c["neighbours"], np.int32), w)


# ---------------------------------------------------------------- hash ----

def test_hash_constants_from_mathematics():
    # SHA-256 K[0..15]: cube roots of the first 16 primes (FIPS 180-4 §4.2.2).
    k = literal.sha256_k(16)
    assert k[0] == 0x428A2F98 and k[15] == 0xC19BF174
    assert len(set(k)) == 16
    # MD5 IV from RFC 1321 §3.3 byte strings.
    assert literal.md5_iv() == [0x67452301, 0xEFCDAB89, 0x98BADCFE, 0x10325476]


def numpy_hash(i, s):
    """Vectorised restatement (written separately from oracle and literal)."""
    i = i.astype(np.uint64)
    s = s.astype(np.uint64)
    m = np.uint64(0xFFFFFFFF)
    h = [np.full(i.shape, v, np.uint64) for v in literal.md5_iv()]
    a, b, c, d = (x.copy() for x in h)
    g = i & m
    for j, kj in enumerate(literal.sha256_k(16)):
        f = ((b & c) | ((~b & m) & d)) & m
        x = (a + f + np.uint64(kj) + g) & m
        r = np.uint64((7, 12, 17, 22)[j % 4])
        rot = ((x << r) | (x >> (np.uint64(32) - r))) & m
        a, b, c, d = d, (b + rot) & m, b, c
        h = [(h[0] + a) & m, (h[1] + b) & m, (h[2] + c) & m, (h[3] + d) & m]
        g = (g * s) & m
    return ((h[0] + h[1] + h[2] + h[3]) & m).astype(np.uint32)


def test_hash_matches_numpy_restatement_1e6():
    rng = np.random.default_rng(1)
    ids = rng.integers(0, 1 << 32, size=1 << 20, dtype=np.uint64).astype(np.uint32)
    seeds = rng.integers(0, 1 << 32, size=1 << 20, dtype=np.uint64).astype(np.uint32)
    ids[:4096] = np.arange(4096, dtype=np.uint32)
    np.testing.assert_array_equal(oracle.select_hash(ids, seeds), numpy_hash(ids, seeds))


def test_hash_matches_committed_vectors():
    with open(os.path.join(GOLD, "hash_vectors.json")) as f:
        vec = np.array(json.load(f)["vectors"], dtype=np.uint64)
    got = oracle.select_hash(vec[:, 0].astype(np.uint32), vec[:, 1].astype(np.uint32))
    np.testing.assert_array_equal(got, vec[:, 2].astype(np.uint32))


def test_hash_special_cases():
    seeds = np.array([0, 1, 2, 12345, 0xFFFFFFFF, 383100999], np.uint32)
    # vertex 0: g == 0 at every step, so the colour does not depend on the seed
    h0 = oracle.select_hash(np.zeros(len(seeds), np.uint32), seeds)
    assert len(set(h0.tolist())) == 1
    # seed 0: g is the id at step 0 and 0 afterwards; seed 1 keeps g == id
    ids = np.arange(1, 1000, dtype=np.uint32)
    assert len(set(oracle.select_hash(ids, 0).tolist())) > 990
    assert not np.array_equal(oracle.select_hash(ids, 1), oracle.select_hash(ids, 0))
    # barrier 0: every vertex red; barrier 2^32-1: blue unless the hash is 2^32-1
    assert all(oracle.colour(i, 77, 0) == 1 for i in range(200))
    assert sum(oracle.colour(i, 77, 0xFFFFFFFF) for i in range(200)) == 0


def test_blue_fraction_tracks_barrier():
    ids = np.arange(1 << 18, dtype=np.uint32)
    for barrier in (0x40000000, 0x88B81733, 0xC0000000):
        frac = float((oracle.select_hash(ids, 987654321) < barrier).mean())
        assert abs(frac - barrier / 2 ** 32) < 0.005


# ------------------------------------------------------------ rounds ----

def test_oracle_matches_golden_tiny_cases():
    data = load_cases()
    seeds = np.array(data["seeds"], np.uint32)
    for c in data["cases"]:
        csr = case_csr(c)
        m, R, _, _ = oracle.match(csr, c["weighted"], c["maximal"], seeds, max_rounds=c["max_rounds"],
                                  barrier=c["barrier"])
        assert R == c["rounds"], c["spec"]
        assert m.tolist() == c["match"], (c["spec"], c["weighted"], c["maximal"])


def test_literal_statement_matches_oracle_on_random_tiny_graphs():
    rng = random.Random(5)
    seeds = reference_seeds(255)
    for trial in range(40):
        n = rng.randrange(2, 14)
        edges = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < 0.4]
        w = [rng.choice([0.25, 0.5, 0.75, 1.0]) for _ in edges]
        csr = CSR.from_edges(n, edges, w, sort_rows=trial % 2 == 0)
        rows = [csr.neighbours[a:b].tolist() for a, b in csr.ranges]
        wts = [csr.weights[a:b].tolist() for a, b in csr.ranges]
        for weighted, maximal in itertools.product((False, True), (False, True)):
            exp, R = literal.run(n, rows, wts, weighted, maximal, seeds)
            got, R2, _, _ = oracle.match(csr, weighted, maximal, seeds)
            assert (got.tolist(), R2) == (exp, R)


def seed_where(pred, limit=1 << 16):
    for s in range(1, limit):
        if pred(s):
            return s
    raise AssertionError("no seed found")


def test_single_edge_closed_form():
    csr = CSR.from_edges(2, [(0, 1)])
    for s in range(1, 60):
        c0, c1 = oracle.colour(0, s, 0x88B81733), oracle.colour(1, s, 0x88B81733)
        m, R, _, _ = oracle.match(csr, True, False, [s] * 20, max_rounds=1)
        if c0 != c1:
            assert m.tolist() == [4, 4]
        else:
            assert m.tolist() == [c0, c1]


def test_red_centre_star_heaviest_blue_leaf_wins_first_on_ties():
    # centre 1, leaves 2..6; leaves 3 and 5 tie for the maximum.  (Vertex 0
    # has a seed-independent colour, so it stays out of the star.)
    w = [0.25, 0.75, 0.5, 0.75, 0.125]
    csr = CSR.from_edges(7, [(1, k) for k in range(2, 7)], w)
    s = seed_where(lambda s: oracle.colour(1, s, 0x88B81733) == 1
                   and all(oracle.colour(k, s, 0x88B81733) == 0 for k in range(2, 7)))
    m, _, _, _ = oracle.match(csr, True, False, [s], max_rounds=1)
    assert m[1] == 5 and m[3] == 5  # leaf 3 is first of the tied leaves in the centre's row
    assert [m[k] for k in (2, 4, 5, 6)] == [0, 0, 0, 0]
    # random matching: the first blue proposer in CSR order
    m, _, _, _ = oracle.match(csr, False, False, [s], max_rounds=1)
    assert m[1] == 5 and m[2] == 5


def test_blue_centre_proposes_to_heaviest_red_first_on_ties():
    w = [0.5, 0.875, 0.25, 0.875]
    csr = CSR.from_edges(6, [(1, k) for k in range(2, 6)], w)
    s = seed_where(lambda s: oracle.colour(1, s, 0x88B81733) == 0
                   and all(oracle.colour(k, s, 0x88B81733) == 1 for k in range(2, 6)))
    m, _, _, _ = oracle.match(csr, True, False, [s], max_rounds=1)
    assert m[1] == 5 and m[3] == 5 and m[5] == 1


def test_dead_semantics_only_blue_vertices_die():
    # isolated vertices 0 and 1: a vertex dies in the first round it is blue
    # and stays red (active) otherwise
    csr = CSR.from_offsets([0, 0, 0], np.zeros(0, np.int32))
    s_red = seed_where(lambda s: oracle.colour(1, s, 0x88B81733) == 1)
    m, R, _, _ = oracle.match(csr, True, False, [s_red], max_rounds=1)
    assert m[1] == 1
    s_blue = seed_where(lambda s: oracle.colour(1, s, 0x88B81733) == 0)
    m, R, _, _ = oracle.match(csr, True, False, [s_blue], max_rounds=1)
    assert m[1] == 2
    # vertex 0 is seed-independent: a barrier just above its hash makes it blue
    b = int(oracle.select_hash(np.zeros(1, np.uint32), 1)[0]) + 1
    m, R, _, _ = oracle.match(CSR.from_offsets([0, 0], np.zeros(0, np.int32)), False, True, [5, 6], barrier=b)
    assert m.tolist() == [2] and R == 2
    # a blue vertex whose neighbours are all matched dies; a red one never does
    path = CSR.from_edges(3, [(0, 1), (1, 2)])
    for s in range(1, 200):
        m, R, _, _ = oracle.match(path, False, True, reference_seeds(255, srand_seed=s))
        assert oracle.test_matching(path, m)[0]
        assert sorted(m.tolist()).count(2) <= 1 and all(x in (2, 4, 5) for x in m.tolist())


def test_unit_weights_weighted_equals_random():
    seeds = reference_seeds(255)
    for seed in range(6):
        rng = np.random.default_rng(seed)
        n = 300
        edges = {(int(a), int(b)) for a, b in rng.integers(0, n, size=(900, 2)) if a != b}
        edges = [(min(e), max(e)) for e in edges]
        csr = CSR.from_edges(n, edges, sort_rows=bool(seed % 2))
        csr.weights = np.ones(csr.nnz, np.float32)
        for maximal in (False, True):
            a = oracle.match(csr, False, maximal, seeds)
            b = oracle.match(csr, True, maximal, seeds)
            assert a[0].tolist() == b[0].tolist() and a[1] == b[1]


def rgg_small(n=4000, seed=3):
    from mmamd import gen_rgg
    return gen_rgg(n, seed=seed, weighted=True)


def test_invariants_valid_and_maximal_at_convergence():
    csr = rgg_small()
    seeds = reference_seeds(255)
    for weighted in (False, True):
        m, R, act, mat = oracle.match(csr, weighted, True, seeds)
        ok, under, multi, invalid, matchable = oracle.test_matching(csr, m)
        assert ok and (under, multi, invalid) == (0, 0, 0)
        assert R < 255 and matchable == 0 and act[-1] == 0
        assert np.all(np.diff(mat) >= 0)
        assert mat[-1] == int((m >= 4).sum())


def test_serial_equals_openmp():
    csr = rgg_small(20000, 9)
    seeds = reference_seeds(255)
    for weighted in (False, True):
        a = oracle.match(csr, weighted, True, seeds, nthreads=1)
        b = oracle.match(csr, weighted, True, seeds, nthreads=8)
        assert a[0].tolist() == b[0].tolist() and a[1:2] == b[1:2]
        np.testing.assert_array_equal(a[2], b[2])


def test_not_enough_seeds_is_an_error():
    csr = rgg_small(500)
    with pytest.raises(ValueError):
        oracle.match(csr, True, False, [1, 2, 3])


# ------------------------------------------------- metrics / baselines ----

def path4():
    # 0 -1.0- 1 -2.0- 2 -3.0- 3
    return CSR.from_edges(4, [(0, 1), (1, 2), (2, 3)], [1.0, 2.0, 3.0])


def test_get_weight_known_answer():
    csr = path4()
    assert oracle.get_weight(csr, np.array([4, 4, 6, 6], np.int32)) == (4.0, 4)
    assert oracle.get_weight(csr, np.array([0, 5, 5, 1], np.int32)) == (2.0, 2)
    assert oracle.get_weight(csr, np.array([0, 1, 2, 0], np.int32)) == (0.0, 0)


def test_test_matching_known_answers():
    csr = path4()
    assert oracle.test_matching(csr, np.array([4, 4, 6, 6], np.int32)) == (True, 0, 0, 0, 0)
    # 0 and 2 share an id but are not adjacent
    assert oracle.test_matching(csr, np.array([4, 1, 4, 1], np.int32))[:4] == (False, 0, 0, 1)
    # three vertices share one id
    assert oracle.test_matching(csr, np.array([5, 5, 5, 0], np.int32))[:4] == (False, 0, 1, 0)
    # a lone matched id
    assert oracle.test_matching(csr, np.array([7, 0, 0, 0], np.int32))[:4] == (False, 1, 0, 0)
    # unmatched 2-3 still matchable
    assert oracle.test_matching(csr, np.array([4, 4, 0, 0], np.int32)) == (True, 0, 0, 0, 2)


def test_test_matching_walk_resumes_after_the_vertex_it_compared():
    # testMatching (src/matchcpu.cu:166-213) takes i1 and i2 from the sorted
    # order; when their values differ it counts i1 as matched to itself and
    # resumes AFTER i2.  Path 0-1-2-3-4 with values [4, 5, 5, 7, 7]: steps
    # (v0, v1) under, (v2, v3) under, (v4, end) under -- three, although only
    # vertex 0 is alone.  Hand-traced from the reference's loop.
    p5 = CSR.from_edges(5, [(0, 1), (1, 2), (2, 3), (3, 4)])
    assert oracle.test_matching(p5, np.array([4, 5, 5, 7, 7], np.int32))[:4] == (False, 3, 0, 0)
    # the walk enters the pair {v3, v4} at v4 only after a lone value: [4, 6, 6, 6] on path4
    # -> (v0, v1) under; then (v2, v3) share 6 and are adjacent: valid pair, no multi
    assert oracle.test_matching(path4(), np.array([4, 6, 6, 6], np.int32))[:4] == (False, 1, 0, 0)
    # [5, 5, 5, 9]: (v0, v1) adjacent pair, v2 a third member (multi), (v3, end) under
    assert oracle.test_matching(path4(), np.array([5, 5, 5, 9], np.int32))[:4] == (False, 1, 1, 0)


def brute_max_weight(n, edges, w):
    best = 0.0
    for k in range(len(edges) + 1):
        for sub in itertools.combinations(range(len(edges)), k):
            used = set()
            ok = True
            for e in sub:
                a, b = edges[e]
                if a in used or b in used:
                    ok = False
                    break
                used |= {a, b}
            if ok:
                best = max(best, sum(w[e] for e in sub))
    return best


def test_cpu_baselines_known_answers_and_half_bound():
    csr = path4()
    assert oracle.cpu_weighted(csr).tolist() == [4, 4, 6, 6]  # 0 takes 1 (only choice), 2 takes 3
    assert oracle.cpu_weighted_edge(csr).tolist() == [4, 4, 6, 6]  # heaviest edge 2-3 first, then 0-1
    assert oracle.cpu_random(csr).tolist() == [4, 4, 6, 6]
    rng = random.Random(11)
    for _ in range(25):
        n = rng.randrange(3, 8)
        edges = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < 0.6]
        if not edges:
            continue
        w = [float(rng.randrange(1, 9)) for _ in edges]
        csr = CSR.from_edges(n, edges, w)
        opt = brute_max_weight(n, edges, w)
        for fn in (oracle.cpu_weighted, oracle.cpu_weighted_edge, oracle.cpu_random):
            m = fn(csr)
            assert oracle.test_matching(csr, m)[0]
            assert oracle.test_matching(csr, m)[4] == 0  # greedy results are maximal
        # heaviest-edge-first greedy is the textbook 1/2-approximation
        assert oracle.get_weight(csr, oracle.cpu_weighted_edge(csr))[0] >= opt / 2 - 1e-9


# --------------------------------------------------------------- seeds ----

def glibc_type3(seed, count):
    """glibc random() TYPE_3 (x^31 + x^3 + 1), as published by P. Selinger,
    'The GLIBC pseudo-random number generator' (2007)."""
    r = [0] * (344 + count)
    r[0] = seed if seed != 0 else 1
    for i in range(1, 31):
        r[i] = (16807 * r[i - 1]) % 2147483647
        if r[i] < 0:
            r[i] += 2147483647
    for i in range(31, 34):
        r[i] = r[i - 31]
    for i in range(34, 344 + count):
        r[i] = (r[i - 31] + r[i - 3]) & 0xFFFFFFFF
    return [r[k + 344] >> 1 for k in range(count)]


def test_reference_seed_stream_is_glibc_rand():
    got = reference_seeds(300).tolist()
    assert got == glibc_type3(12345, 300)
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(ctypes.c_uint(1))
    libc.rand.restype = ctypes.c_int
    assert libc.rand() == glibc_type3(1, 1)[0] == 1804289383
