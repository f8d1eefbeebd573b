"""GPU checks of the autotuner and of the batched path at the batched
workload's graph mix: tuning never changes a result, and every graph of a
batch matches exactly as it does alone."""
import sys

import numpy as np
import pytest

import mmamd
import oracle
import workloads
from mmamd import MM_WEIGHTED, Batch, Graph, reference_seeds

pytestmark = pytest.mark.gpu

SEEDS = reference_seeds(255)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def tiles(t):
    return dict(lanes=t["lanes_per_vertex"], block_size=t["block_size"], hub_degree=t["hub_degree"],
                unroll=t["entries_per_thread"])


@pytest.mark.parametrize("name,maximal", [("rgg16", False), ("rgg16", True), ("rmat14", True)])
def test_autotuner_candidates_agree_and_pick_is_exact(name, maximal):
    csr = mmamd.gen_rgg(1 << 16, seed=7, weighted=True) if name == "rgg16" else \
        mmamd.gen_rmat(14, seed=8, weighted=True)
    g = Graph(csr)
    t = g.autotune(kind=MM_WEIGHTED, maximal=maximal, max_rounds=0 if maximal else 5, seeds=SEEDS, reps=2)
    say(name, maximal, t)
    assert t["mismatches"] == 0
    assert t["candidates"] >= 15
    assert t["best_ms"] <= t["default_ms"]
    res = g.match(kind=MM_WEIGHTED, maximal=maximal, max_rounds=0 if maximal else 5, seeds=SEEDS, **tiles(t))
    exp, R, _, _ = oracle.match(csr, True, maximal, SEEDS, max_rounds=0 if maximal else 5, nthreads=0)
    assert res.rounds == R
    np.testing.assert_array_equal(res.match, exp)


def test_batch_of_grids_and_rggs_equals_single_graph_path():
    specs = workloads.batch_sizes(192, 8, 12, seed=99)
    graphs = [workloads.batch_graph(*s) for s in specs]
    b = Batch(graphs)
    res, per_graph, rounds = b.match(kind=MM_WEIGHTED, maximal=True, seeds=SEEDS)
    for k, csr in enumerate(graphs):
        single = Graph(csr).match(kind=MM_WEIGHTED, maximal=True, seeds=SEEDS)
        assert single.rounds == rounds[k], (k, specs[k])
        np.testing.assert_array_equal(single.match, per_graph[k], err_msg=str((k, specs[k])))
    assert res.rounds == rounds.max()
    union, vo = workloads.concat(graphs)
    ok, *counts = mmamd.test_matching(union, workloads.batch_match_global(res.match, vo))
    assert ok and counts[-1] == 0
    t = b.autotune(kind=MM_WEIGHTED, maximal=True, seeds=SEEDS, reps=1)
    say("batch autotune", t)
    assert t["mismatches"] == 0
    tuned, _, rounds2 = b.match(kind=MM_WEIGHTED, maximal=True, seeds=SEEDS, **tiles(t))
    np.testing.assert_array_equal(tuned.match, res.match)
    np.testing.assert_array_equal(rounds2, rounds)
