"""GPU parity of graphs with many vertices without neighbours.

When at least 1% of the vertices have no neighbour, the first round's select
settles them at once (their colour in every round depends only on their id
and the seeds: blue means dead, red means they stay) and keeps them out of
the work lists; the per-round counters get their share afterwards.  These
tests compare the whole match array, the round count and the per-round
active and matched counts with the oracle, which runs the rounds literally,
for fixed and maximal runs of every length from 1 round up, both matching
kinds, both response modes, both vertex orders, and a batch.
"""
import sys

import numpy as np
import pytest

import mmamd
import oracle
from mmamd import MM_RANDOM, MM_RESPOND_PULL, MM_RESPOND_PUSH, MM_WEIGHTED, Batch, Graph, reference_seeds

pytestmark = pytest.mark.gpu

SEEDS = reference_seeds(255)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sparse_graphs():
    # R-MAT leaves many ids without edges; a thin RGG leaves about 7% isolated
    return [("rmat13", mmamd.gen_rmat(13, 8, seed=5, weighted=True)),
            ("rgg14_thin", mmamd.gen_rgg(1 << 14, seed=9, weighted=True, radius_scale=0.3))]


def isolated_fraction(csr):
    r = csr.ranges.reshape(-1, 2)
    return float(np.mean(r[:, 1] == r[:, 0]))


def compare(csr, g, weighted, maximal, max_rounds, **kw):
    exp, R, act, mat = oracle.match(csr, weighted, maximal, SEEDS, max_rounds=max_rounds, nthreads=0)
    res = g.match(kind=MM_WEIGHTED if weighted else MM_RANDOM, maximal=maximal, seeds=SEEDS,
                  max_rounds=max_rounds, **kw)
    assert res.rounds == R
    mism = np.flatnonzero(res.match != exp)
    assert mism.size == 0, f"{mism.size} mismatches, first at {mism[:5]}"
    np.testing.assert_array_equal(res.active[:R], act)
    np.testing.assert_array_equal(res.matched[:R], mat)


@pytest.mark.parametrize("keep", [False, True])
def test_isolated_vertices_settled_by_the_select_match_the_oracle(keep):
    for name, csr in sparse_graphs():
        assert isolated_fraction(csr) >= 0.01, name
        g = Graph(csr, keep_order=keep)
        assert g.info()["iso_settled"] == 1, name
        modes = [MM_RESPOND_PULL] + ([MM_RESPOND_PUSH] if g.info()["push_ok"] else [])
        for weighted in (True, False):
            for respond in modes:
                for maximal, rounds in ((True, 0), (False, 1), (False, 2), (False, 3), (False, 7), (True, 2)):
                    if maximal and rounds:
                        continue
                    for lanes, block, hub in ((8, 1024, 256), (4, 64, 16)):
                        say(name, "keep", keep, "weighted", weighted, "respond", respond, "maximal", maximal,
                            "rounds", rounds, "tiles", lanes, block, hub)
                        compare(csr, g, weighted, maximal, rounds, lanes=lanes, block_size=block,
                                hub_degree=hub, respond=respond)
        g.close()


def test_graphs_without_any_edge():
    """Every vertex without neighbours: the select settles all of them, the
    first round's request and match scan an empty id range, and a maximal run
    ends when the last red vertex has been blue once."""
    for n in (1, 100, 70000):
        for weighted in (True, False):
            w = np.ones(0, np.float32) if weighted else None
            csr = mmamd.CSR.from_offsets(np.zeros(n + 1, np.int64), np.zeros(0, np.int32), w)
            for keep in (False, True):
                g = Graph(csr, keep_order=keep)
                modes = [MM_RESPOND_PULL] + ([MM_RESPOND_PUSH] if g.info()["push_ok"] else [])
                for respond in modes:
                    for maximal, rounds in ((True, 0), (False, 1), (False, 4)):
                        say("edgeless n", n, "weighted", weighted, "keep", keep, "respond", respond,
                            "maximal", maximal, "rounds", rounds)
                        compare(csr, g, weighted, maximal, rounds, respond=respond)
                g.close()


def test_threshold_of_settling_vertices_without_neighbours():
    """The select settles vertices without neighbours when they are at least
    1% of the graph: exactly 1% takes that path, one vertex fewer does not, and
    both give the oracle's result."""
    base = mmamd.gen_rgg(9900, seed=41, weighted=True, radius_scale=1.2)
    r = base.ranges
    assert np.all(r[:, 1] > r[:, 0])
    for extra, settled in ((100, 1), (99, 0)):
        n = base.n + extra
        offsets = np.concatenate([r[:, 0], [r[-1, 1]] * (extra + 1)]).astype(np.int64)
        csr = mmamd.CSR.from_offsets(offsets, base.neighbours, base.weights)
        g = Graph(csr)
        assert g.info()["iso_settled"] == (1 if (n - base.n) * 100 >= n else 0) == settled, (extra, n)
        for maximal, rounds in ((True, 0), (False, 3)):
            say("threshold extra", extra, "maximal", maximal)
            compare(csr, g, True, maximal, rounds)
        g.close()


def test_graphs_without_many_isolated_vertices_keep_them_in_the_rounds():
    csr = mmamd.gen_rgg(1 << 14, seed=3, weighted=True)
    assert isolated_fraction(csr) < 0.01
    g = Graph(csr)
    assert g.info()["iso_settled"] == 0
    compare(csr, g, True, True, 0)
    g.close()


def test_batch_of_sparse_graphs_per_graph_results_and_rounds():
    say("batch of sparse graphs")
    graphs = [mmamd.gen_rmat(s, 8, seed=10 + s, weighted=True) for s in (6, 8, 10, 11)]
    graphs += [mmamd.gen_rgg(n, seed=n, weighted=True, radius_scale=0.3) for n in (100, 700, 3000)]
    graphs.append(mmamd.gen_rgg(0, seed=1))
    b = Batch(graphs)
    assert b.info()["iso_settled"] == 1
    for maximal, rounds in ((True, 0), (False, 3)):
        res, per, rk = b.match(kind=MM_WEIGHTED, maximal=maximal, seeds=SEEDS, max_rounds=rounds)
        for k, csr in enumerate(graphs):
            exp, R, _, _ = oracle.match(csr, True, maximal, SEEDS, max_rounds=rounds)
            assert per[k].tolist() == exp.tolist(), k
            assert rk[k] == R, (k, rk[k], R)
        assert res.rounds == max(rk)
