"""The device result check (mm_graph_check / mm_batch_check: getWeight and
testMatching, src/matchcpu.cu:64-246) against the CPU oracle's restatements,
on the matchers' results and on planted-invalid arrays, singly and per graph
of a batch.

The weights of the generated graphs are multiples of 2^-24 below 1, so every
sum of them here is exact in double in any order: weights must be equal."""
import sys

import numpy as np
import pytest

import mmamd
import oracle
import workloads
from check_util import plant_invalid
from mmamd import MM_WEIGHTED, Batch, Graph, reference_seeds

pytestmark = pytest.mark.gpu

SEEDS = reference_seeds(255)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_check(csr, m):
    ok, under, multi, invalid, matchable = oracle.test_matching(csr, m)
    w, size = oracle.get_weight(csr, m)
    return dict(under=under, multi=multi, invalid=invalid, matchable=matchable, size=size, weight=w, valid=int(ok))


def assert_check(dev, exp, what):
    for k in ("under", "multi", "invalid", "matchable", "size", "valid", "weight"):
        assert dev[k] == exp[k], (what, k, dev, exp)


@pytest.mark.parametrize("keep", [False, True])
def test_device_check_equals_oracle(keep):
    csr = mmamd.gen_rgg(1 << 15, seed=17, weighted=True)
    g = Graph(csr, keep_order=keep)
    rng = np.random.default_rng(3)
    for maximal in (False, True):
        say("keep", keep, "maximal", maximal)
        res = g.match(kind=MM_WEIGHTED, maximal=maximal, max_rounds=0 if maximal else 3, seeds=SEEDS)
        assert_check(g.check(), oracle_check(csr, res.match), "result")
        bad = plant_invalid(res.match, csr, rng)
        exp = oracle_check(csr, bad)
        assert exp["under"] and exp["multi"] and exp["invalid"] and not exp["valid"]
        g.set_match(bad)
        assert_check(g.check(), exp, "planted")
        np.testing.assert_array_equal(g.copy_match(), bad)
    g.close()


def test_batch_device_check_per_graph_equals_oracle():
    specs = workloads.batch_sizes(64, 8, 12, seed=5)
    graphs = [workloads.batch_graph(*s) for s in specs]
    b = Batch(graphs)
    res, per_graph, _ = b.match(kind=MM_WEIGHTED, maximal=True, seeds=SEEDS)
    rng = np.random.default_rng(11)
    arrays = [("result", [np.array(p) for p in per_graph])]
    planted = [np.array(p) for p in per_graph]
    for k in range(0, len(graphs), 3):
        planted[k] = plant_invalid(planted[k], graphs[k], rng)
    for k in range(1, len(graphs) - 1, 5):
        # a graph-local value just past the graph's end: the next graph's first
        # id in the concatenated numbering, still a group of this graph alone
        a, b2 = graphs[k].ranges[0]
        planted[k][0] = 4 + graphs[k].n
        if b2 > a:
            planted[k][graphs[k].neighbours[a]] = 4 + graphs[k].n
    arrays.append(("planted", planted))
    for what, arrs in arrays:
        say("batch", what)
        b_set = np.concatenate(arrs).astype(np.int32)
        b._set_match(b_set)
        tot, gw, gs, gv = b.check()
        exp = [oracle_check(csr, arrs[k]) for k, csr in enumerate(graphs)]
        for k in range(len(graphs)):
            assert gs[k] == exp[k]["size"] and gw[k] == exp[k]["weight"] and gv[k] == exp[k]["valid"], (what, k)
        sums = {key: sum(e[key] for e in exp) for key in ("under", "multi", "invalid", "matchable", "size")}
        for key, val in sums.items():
            assert tot[key] == val, (what, key, tot, sums)
        assert tot["weight"] == sum(e["weight"] for e in exp)
        assert tot["valid"] == int(all(e["valid"] for e in exp))
    b.close()


@pytest.mark.parametrize("keep", [False, True])
def test_device_check_weight_is_the_csr_first_entry(keep):
    """getWeight takes the weight at the first CSR position whose neighbour
    shares the value: a repeated neighbour's lighter copy when it comes first,
    a self-loop when it precedes the partner."""
    rows = [[(1, 0.25), (1, 0.75), (2, 0.5)],          # 0: neighbour 1 twice, lighter copy first
            [(0, 0.75), (0, 0.25), (4, 0.5)],          # 1: heavier copy first
            [(2, 0.125), (3, 0.5), (0, 0.5)],          # 2: self-loop ahead of partner 3
            [(2, 0.5), (5, 0.375)],
            [(1, 0.5), (5, 0.625)],
            [(3, 0.375), (4, 0.625)]]
    offsets = np.cumsum([0] + [len(r) for r in rows])
    nb = [u for r in rows for u, _ in r]
    w = [x for r in rows for _, x in r]
    csr = mmamd.CSR.from_offsets(offsets, nb, w)
    g = Graph(csr, keep_order=keep)
    arrays = [np.array([4, 4, 6, 6, 8, 8], np.int32),    # {0,1} {2,3} {4,5}
              np.array([4, 4, 6, 6, 0, 1], np.int32),
              np.array([4, 4, 4, 6, 6, 2], np.int32),    # a group of three, a pair that is no edge
              np.array([4, 4, 9, 9, 30, 30], np.int32)]  # a value past the end
    for m in arrays:
        g.set_match(m)
        exp = oracle_check(csr, m)
        assert_check(g.check(), exp, ("keep", keep, m.tolist()))
    g.close()
