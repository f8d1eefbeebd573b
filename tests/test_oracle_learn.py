"""CPU checks of the learned-selection oracle (oracle/mm_oracle_learn.c) and of
the float64 network oracle in oracle/oracle.py.  The learned selection has no
reference counterpart; these checks pin the oracle to mathematics (sigmoid,
finite differences), to the plain oracle (params = None) and to closed forms."""
import numpy as np
import pytest

import mmamd
import oracle
from mmamd import CSR, reference_seeds

SEEDS = reference_seeds(255)


def random_params(seed, scale=1.0):
    return (mmamd.mlp_init(seed) * np.float32(scale)).astype(np.float32)


def bit_network():
    """z = 60 * x7 - 30: blue exactly when the vertex's hash is below that of
    its heaviest unmatched neighbour."""
    p = np.zeros(oracle.NP, np.float32)
    p[0 * 8 + 7] = 1.0           # W1[0][7]: h1[0] = x7
    p[288 + 0 * 32 + 0] = 1.0    # W2[0][0]: h2[0] = h1[0]
    p[1344 + 0] = 60.0           # W3[0]
    p[1376] = -30.0              # b3
    return p


def const_network(z):
    p = np.zeros(oracle.NP, np.float32)
    p[1376] = z
    return p


def test_blue_threshold_is_65536_sigmoid():
    z = np.linspace(-40.0, 40.0, 400001).astype(np.float32)
    T = oracle.blue_threshold(z).astype(np.int64)
    ref = np.floor(65536.0 / (1.0 + np.exp(-z.astype(np.float64))))
    assert np.max(np.abs(T - ref)) <= 1
    assert np.all(np.diff(T) >= 0)
    assert oracle.blue_threshold(np.float32([0.0]))[0] == 32768
    assert oracle.blue_threshold(np.float32([30.0]))[0] == 65536
    assert oracle.blue_threshold(np.float32([-30.0]))[0] == 0
    assert oracle.blue_threshold(np.float32([np.nan]))[0] == 0


def test_mlp_z_matches_float64_forward():
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 2, (5000, 8)).astype(np.float32)
    for s in range(3):
        p = random_params(s)
        z32 = oracle.mlp_z(p, x).astype(np.float64)
        z64, _ = oracle.mlp_forward64(p, x)
        np.testing.assert_allclose(z32, z64, rtol=1e-5, atol=1e-5)


def test_float64_gradient_matches_finite_differences():
    rng = np.random.default_rng(4)
    x = rng.uniform(0, 1.5, (64, 8))
    y = (rng.uniform(size=64) < 0.5).astype(np.float64)
    w = rng.uniform(0.1, 1.0, 64)
    p = random_params(5).astype(np.float64)
    loss, g = oracle.mlp_loss_grad64(p, x, y, w)
    h = 1e-6
    for e in list(range(0, 1377, 37)) + [255, 256, 287, 288, 1311, 1312, 1343, 1344, 1375, 1376]:
        q = p.copy()
        q[e] += h
        lp = oracle.mlp_loss_grad64(q, x, y, w)[0]
        q[e] -= 2 * h
        lm = oracle.mlp_loss_grad64(q, x, y, w)[0]
        assert abs((lp - lm) / (2 * h) - g[e]) <= 1e-6 + 1e-5 * abs(g[e]), e


@pytest.mark.parametrize("weighted,maximal", [(False, False), (False, True), (True, False), (True, True)])
def test_params_none_is_the_plain_oracle(weighted, maximal):
    for csr in (mmamd.gen_rgg(3000, seed=5, weighted=True), mmamd.gen_rmat(11, seed=9, weighted=True),
                mmamd.gen_grid(40, 30, seed=2, weighted=True)):
        exp, R, act, mat = oracle.match(csr, weighted, maximal, SEEDS)
        got, R2, act2, mat2, _ = oracle.match_learned(csr, weighted, maximal, SEEDS)
        assert R == R2
        np.testing.assert_array_equal(exp, got)
        np.testing.assert_array_equal(act, act2)
        np.testing.assert_array_equal(mat, mat2)


def test_disjoint_edges_match_in_one_round_with_the_hash_order_network():
    k = 500
    edges = np.array([[2 * i, 2 * i + 1] for i in range(k)])
    w = np.linspace(0.1, 1.0, k).astype(np.float32)
    csr = CSR.from_edges(2 * k, edges, w)
    m, R, act, mat, _ = oracle.match_learned(csr, True, True, SEEDS, params=bit_network())
    assert R == 2 and mat[0] == 2 * k and act[1] == 0
    assert np.array_equal(m, 4 + (np.arange(2 * k) & ~1))
    # the plain colouring needs several rounds on the same graph
    assert oracle.match(csr, True, True, SEEDS)[1] > 2


def test_all_red_network_matches_nothing_and_retires_only_isolated_vertices():
    edges = np.array([[0, 1], [1, 2], [3, 4]])
    csr = CSR.from_edges(6, edges, np.float32([0.5, 0.7, 0.9]))   # vertex 5 is isolated
    m, R, act, mat, _ = oracle.match_learned(csr, True, False, SEEDS, params=const_network(-30.0), max_rounds=3)
    assert R == 3 and list(mat) == [0, 0, 0]
    assert list(m) == [1, 1, 1, 1, 1, 2]
    m, R, act, mat, _ = oracle.match_learned(csr, True, False, SEEDS, params=const_network(30.0), max_rounds=3)
    assert list(mat) == [0, 0, 0] and list(m) == [0, 0, 0, 0, 0, 2]


def test_plain_traces_record_matched_vertices_with_colour_and_edge_weight():
    csr = mmamd.gen_rgg(2000, seed=11, weighted=True)
    m, R, act, mat, s = oracle.match_learned(csr, True, True, SEEDS, trace=True)
    assert s["offered"] == mat[-1] == len(s["y"])
    r = csr.ranges
    for i in range(len(s["v"])):
        v, rd = int(s["v"][i]), int(s["round"][i])
        blue = oracle.select_hash(np.uint32([v]), SEEDS[rd - 1])[0] < 0x88B81733
        assert s["y"][i] == (1.0 if blue else 0.0)
        partner = m[v] - 4 if m[v] - 4 != v else None
        if partner is None:   # v is the smaller id: find the neighbour holding the same match value
            nb = csr.neighbours[r[v, 0]:r[v, 1]]
            partner = int(nb[np.flatnonzero(m[nb] == m[v])[0]])
        j = r[v, 0] + np.flatnonzero(csr.neighbours[r[v, 0]:r[v, 1]] == partner)[0]
        assert s["w"][i] == csr.weights[j]
        assert s["x"][i, 6] == (rd - 1) / 32.0
        assert s["x"][i, 0] == min(r[v, 1] - r[v, 0], 1024) / 64.0
    order = np.lexsort((s["v"], s["round"]))
    assert np.array_equal(order, np.arange(len(order)))
    _, _, _, _, s0 = oracle.match_learned(csr, True, True, SEEDS, trace=True, trace_keep=0)
    assert s0["offered"] == 0


def test_learned_results_are_valid_matchings():
    csr = mmamd.gen_rgg(4000, seed=21, weighted=True)
    for s in range(3):
        m, R, act, mat, _ = oracle.match_learned(csr, True, True, SEEDS, params=random_params(s))
        ok, *counts = oracle.test_matching(csr, m)
        assert ok and counts[-1] == 0, counts   # valid and maximal at convergence
        assert np.all(np.diff(mat) >= 0)
