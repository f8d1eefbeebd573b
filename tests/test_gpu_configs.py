"""GPU parity at the benchmark configurations' full sizes, through the C ABI,
against the CPU oracle:

  - R-MAT scale 24 (n = 2^24, 521 M row entries, hub rows of 10^4-10^5
    entries, 64-bit proposal keys), weighted maximal rounds;
  - the 4096-graph batch (n log-uniform in [2^8, 2^16], grids and RGGs):
    every graph's result and round count against the oracle run on that graph
    alone, and the device check per graph against the oracle's getWeight /
    testMatching;
  - the learned selection on the 2^22-vertex RGG with the committed network:
    match array, rounds and per-round counts against the oracle's learned
    rounds, and the learned result's weight and rounds against the plain one.

Integer/index work: the bar is bit-exact equality."""
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import learn
import mmamd
import oracle
import workloads
from mmamd import MM_WEIGHTED, Batch, Graph, reference_seeds

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

SEEDS = reference_seeds(255)
HOST_THREADS = min(16, os.cpu_count() or 1)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_check(csr, m):
    """getWeight / testMatching of the oracle, in the device check's layout.
    The weights are multiples of 2^-24 below 1, so the sums are exact in any
    order and are compared for equality."""
    ok, under, multi, invalid, matchable = oracle.test_matching(csr, m)
    w, size = oracle.get_weight(csr, m)
    return dict(under=under, multi=multi, invalid=invalid, matchable=matchable, size=size, weight=w, valid=int(ok))


def test_rmat24_weighted_maximal_bitexact():
    t0 = time.time()
    csr = workloads.rmat(24)
    say(f"rmat24 generated: n={csr.n} nnz={csr.nnz} ({time.time() - t0:.1f} s)")
    g = Graph(csr)
    say(f"rmat24 uploaded ({time.time() - t0:.1f} s); key bits {g.info()['key_bits']}, "
        f"max degree {g.info()['max_degree']}")
    res = g.match(kind=MM_WEIGHTED, maximal=True, seeds=SEEDS)
    exp, R, act, mat = oracle.match(csr, True, True, SEEDS, nthreads=0)
    say(f"rmat24 rounds gpu {res.rounds} oracle {R} ({time.time() - t0:.1f} s)")
    assert res.rounds == R
    mism = np.flatnonzero(res.match != exp)
    assert mism.size == 0, f"{mism.size} mismatches, first at {mism[:5]}"
    np.testing.assert_array_equal(res.active, act)
    np.testing.assert_array_equal(res.matched, mat)
    ok, under, multi, invalid, matchable = oracle.test_matching(csr, res.match)
    assert ok and matchable == 0
    assert g.check() == oracle_check(csr, res.match)
    # the benchmark step: the rounds up to 99% of the final matched count
    r99 = int(np.argmax(mat >= 0.99 * mat[-1])) + 1
    fixed = g.match(kind=MM_WEIGHTED, maximal=False, max_rounds=r99, seeds=SEEDS)
    exp99 = oracle.match(csr, True, False, SEEDS, max_rounds=r99, nthreads=0)[0]
    np.testing.assert_array_equal(fixed.match, exp99)
    assert g.check() == oracle_check(csr, exp99)
    g.close()


def test_batch4096_every_graph_equals_the_oracle():
    t0 = time.time()
    specs = workloads.batch_sizes(4096)
    graphs = [workloads.batch_graph(*s) for s in specs]
    say(f"batch generated: {sum(c.n for c in graphs)} vertices ({time.time() - t0:.1f} s)")
    b = Batch(graphs)
    say(f"batch uploaded ({time.time() - t0:.1f} s)")
    res, per_graph, rounds = b.match(kind=MM_WEIGHTED, maximal=True, seeds=SEEDS)
    say(f"batch matched in {res.rounds} rounds ({time.time() - t0:.1f} s)")

    def one(k):
        csr = graphs[k]
        exp, R, _, _ = oracle.match(csr, True, True, SEEDS, nthreads=1)
        ok, under, multi, invalid, matchable = oracle.test_matching(csr, exp)
        w, size = oracle.get_weight(csr, exp)
        return k, np.array_equal(exp, per_graph[k]), R, (w, size, ok and matchable == 0)

    with ThreadPoolExecutor(HOST_THREADS) as pool:
        results = list(pool.map(one, range(len(graphs))))
    say(f"oracle done ({time.time() - t0:.1f} s)")
    bad = [k for k, same, R, _ in results if not same or R != rounds[k]]
    assert not bad, f"{len(bad)} graphs differ from the oracle, first {bad[:5]}: {[specs[k] for k in bad[:3]]}"
    assert res.rounds == rounds.max()
    t1 = time.time()
    tot, gw, gs, gv = b.check()
    say(f"device check {tot} in {(time.time() - t1) * 1e3:.1f} ms")
    wrong = [(k, valid, bool(gv[k]), int(gs[k]), size, float(gw[k]), w)
             for k, _, _, (w, size, valid) in results
             if not (valid and gv[k] and gs[k] == size and gw[k] == w)]
    assert not wrong, f"{len(wrong)} graphs' device checks differ (graph, oracle valid, device valid, device size, " \
                      f"oracle size, device weight, oracle weight): {wrong[:6]}"
    assert tot["valid"] and tot["matchable"] == 0
    assert tot["size"] == sum(r[3][1] for r in results)
    assert tot["weight"] == sum(r[3][0] for r in results)  # weights are multiples of 2^-24: exact sums
    b.close()


def test_learned_rgg22_equals_the_oracle_and_beats_plain():
    t0 = time.time()
    csr = workloads.rgg(22)
    params = learn.load_params()
    g = Graph(csr)
    res = g.match_learned(params, kind=MM_WEIGHTED, maximal=True, seeds=SEEDS)
    say(f"rgg22 learned on the GPU: {res.rounds} rounds ({time.time() - t0:.1f} s)")
    exp, R, act, mat, _ = oracle.match_learned(csr, True, True, SEEDS, params=params)
    say(f"oracle learned: {R} rounds ({time.time() - t0:.1f} s)")
    assert res.rounds == R
    mism = np.flatnonzero(res.match != exp)
    assert mism.size == 0, f"{mism.size} mismatches, first at {mism[:5]}"
    np.testing.assert_array_equal(res.active, act)
    np.testing.assert_array_equal(res.matched, mat)
    plain = g.match(kind=MM_WEIGHTED, maximal=True, seeds=SEEDS)
    w_learned, size_learned = oracle.get_weight(csr, res.match)
    w_plain, _ = oracle.get_weight(csr, plain.match)
    say(f"learned weight {w_learned} in {res.rounds} rounds, plain {w_plain} in {plain.rounds}")
    assert oracle.test_matching(csr, res.match)[0]
    assert w_learned >= w_plain and res.rounds < plain.rounds
    g.close()
