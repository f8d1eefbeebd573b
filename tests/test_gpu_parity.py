"""GPU parity: the HIP matcher (through the C ABI) against the golden fixtures
and the CPU oracle.  Integer/index work, so the bar is bit-exact equality of
the whole match array and of the round count.
"""
import json
import os
import sys

import numpy as np
import pytest

import mmamd
import oracle
from mmamd import (CSR, MM_RANDOM, MM_RESPOND_PULL, MM_RESPOND_PUSH, MM_WEIGHTED, Batch, Graph, reference_seeds)

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SEEDS = reference_seeds(255)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def case_csr(c):
    w = np.array(c["weight_bits"], dtype=np.uint32).view(np.float32)
    return CSR.from_offsets(c["offsets"], np.array(c["neighbours"], np.int32), w)


def check_against_oracle(csr, g, weighted, maximal, **kw):
    exp, R, act, mat = oracle.match(csr, weighted, maximal, SEEDS, nthreads=0)
    res = g.match(kind=MM_WEIGHTED if weighted else MM_RANDOM, maximal=maximal, seeds=SEEDS, **kw)
    assert res.rounds == R
    mism = np.flatnonzero(res.match != exp)
    assert mism.size == 0, f"{mism.size} mismatches, first at {mism[:5]}"
    np.testing.assert_array_equal(res.active, act)
    np.testing.assert_array_equal(res.matched, mat)
    return res


def test_device_present():
    assert mmamd.lib().mm_device_count() >= 1


def test_golden_tiny_cases_all_types_and_tiles():
    with open(os.path.join(HERE, "golden", "tiny_cases.json")) as f:
        data = json.load(f)
    seeds = np.array(data["seeds"], np.uint32)
    for idx, c in enumerate(data["cases"]):
        say("case", idx, c["spec"], "weighted", c["weighted"], "maximal", c["maximal"])
        csr = case_csr(c)
        for keep in (False, True):
            g = Graph(csr, keep_order=keep)
            modes = [MM_RESPOND_PULL] + ([MM_RESPOND_PUSH] if g.info()["push_ok"] else [])
            for respond in modes:
                for lanes, block, hub in ((-1, 256, 256), (-1, 64, 3), (-1, 1024, 1), (0, 256, 256), (1, 64, 2),
                                          (4, 256, 256), (16, 256, 3), (64, 128, 1)):
                    res = g.match(kind=MM_WEIGHTED if c["weighted"] else MM_RANDOM, maximal=c["maximal"],
                                  seeds=seeds, max_rounds=c["max_rounds"], barrier=c["barrier"], lanes=lanes,
                                  block_size=block, hub_degree=hub, respond=respond)
                    assert res.rounds == c["rounds"], (idx, keep, respond, lanes)
                    assert res.match.tolist() == c["match"], (idx, keep, respond, lanes, block, hub)
            g.close()


@pytest.mark.parametrize("maximal", [False, True])
def test_rgg16_random_bitexact(maximal):
    """RGG n=2^16, unit weights, random matching (types 6/7), one seed per round."""
    say("rgg16 random maximal", maximal)
    csr = mmamd.gen_rgg(1 << 16, seed=12345, weighted=False)
    g = Graph(csr)
    res = check_against_oracle(csr, g, False, maximal)
    assert oracle.test_matching(csr, res.match)[0]


@pytest.mark.parametrize("maximal,keep", [(False, False), (True, False), (True, True)])
def test_rgg18_weighted_bitexact(maximal, keep):
    say("rgg18 weighted maximal", maximal, "keep_order", keep)
    csr = mmamd.gen_rgg(1 << 18, seed=777, weighted=True)
    g = Graph(csr, keep_order=keep)
    assert g.info()["push_ok"] == 1
    for respond in (MM_RESPOND_PULL, MM_RESPOND_PUSH):
        res = check_against_oracle(csr, g, True, maximal, respond=respond)
        assert res.respond == respond
    assert oracle.test_matching(csr, res.match)[0]


def test_unit_weights_weighted_equals_random_on_gpu():
    say("unit weights identity")
    csr = mmamd.gen_rgg(1 << 15, seed=5, weighted=False)
    g = Graph(csr)
    a = g.match(kind=MM_RANDOM, maximal=True, seeds=SEEDS)
    b = g.match(kind=MM_WEIGHTED, maximal=True, seeds=SEEDS)
    assert a.rounds == b.rounds and np.array_equal(a.match, b.match)


def shuffled(csr, seed):
    """The reference's random_shuffle (src/graph.cu:513-548): rows move with
    their vertex, neighbour ids are relabelled in place, so rows end up
    unsorted."""
    rng = np.random.default_rng(seed)
    n = csr.n
    perm = rng.permutation(n)
    inv = np.empty(n, np.int64)
    inv[perm] = np.arange(n)
    ranges = csr.ranges[perm]
    nb = inv[csr.neighbours].astype(np.int32)
    return CSR(np.ascontiguousarray(ranges, dtype=np.int32), nb, csr.weights)


@pytest.mark.parametrize("keep", [False, True])
def test_shuffled_rows_use_exact_pull_path(keep):
    say("shuffled rgg keep_order", keep)
    csr = shuffled(mmamd.gen_rgg(1 << 16, seed=31, weighted=True), 4)
    g = Graph(csr, keep_order=keep)
    assert g.info()["push_ok"] == 0
    for weighted in (False, True):
        for lanes in (-1, 0, 2):
            res = check_against_oracle(csr, g, weighted, True, lanes=lanes)
            assert res.respond == MM_RESPOND_PULL


def test_tile_choices_do_not_change_results():
    say("tile determinism")
    csr = mmamd.gen_rgg(1 << 15, seed=8, weighted=True)
    g = Graph(csr)
    ref = g.match(kind=MM_WEIGHTED, maximal=True, seeds=SEEDS)
    for lanes in (-1, 0, 1, 2, 4, 8, 16, 32, 64):
        for block in (64, 256, 1024):
            for hub in (4, 12, 256):
                res = g.match(kind=MM_WEIGHTED, maximal=True, seeds=SEEDS, lanes=lanes, block_size=block,
                              hub_degree=hub)
                assert res.rounds == ref.rounds and np.array_equal(res.match, ref.match), (lanes, block, hub)


@pytest.mark.parametrize("packed", [False, True])
def test_row_layouts_match_the_oracle(packed):
    """Line-aligned device rows (the default) and back-to-back rows
    (MM_CREATE_PACKED_ROWS) give the oracle's result; low-degree grids and
    R-MAT hubs exercise both padding cases, tiles and lane scans both."""
    say("row layout packed", packed)
    for csr in (mmamd.gen_grid(90, 70, seed=6, weighted=True), mmamd.gen_rmat(13, seed=7, weighted=True),
                mmamd.gen_rgg(1 << 14, seed=8, weighted=True)):
        g = Graph(csr, packed_rows=packed)
        for weighted in (False, True):
            for lanes in (-1, 2, 8):
                check_against_oracle(csr, g, weighted, True, lanes=lanes, hub_degree=32)
        g.close()


def test_round_one_without_head_copy():
    """MM_CREATE_NO_HEADS: weighted PUSH rounds then start from work lists
    (the round-1 select files every vertex) instead of the buckets' id ranges
    and the head copy; the result is the oracle's either way."""
    say("no head copy")
    for csr in (mmamd.gen_rgg(1 << 15, seed=31, weighted=True), mmamd.gen_rmat(13, seed=32, weighted=True)):
        for heads in (True, False):
            g = Graph(csr, heads=heads)
            for lanes in (2, 8):
                check_against_oracle(csr, g, True, True, lanes=lanes, hub_degree=64)
                # a fixed 3-round run ends with vertices still active: their colours are part of the result
                exp = oracle.match(csr, True, False, SEEDS, max_rounds=3)[0]
                res = g.match(kind=MM_WEIGHTED, maximal=False, max_rounds=3, seeds=SEEDS, lanes=lanes, hub_degree=64)
                np.testing.assert_array_equal(res.match, exp)
            g.close()


@pytest.mark.parametrize("hub,keep", [(16, False), (256, False), (16, True)])
def test_rmat_hub_path_bitexact(hub, keep):
    say("rmat scale 15 hub", hub, "keep_order", keep)
    csr = mmamd.gen_rmat(15, 16, seed=99, weighted=True)
    g = Graph(csr, keep_order=keep)
    assert g.info()["max_degree"] > 256
    for weighted in (False, True):
        for respond in (MM_RESPOND_PULL, MM_RESPOND_PUSH):
            for lanes in (-1, 0, 32):
                check_against_oracle(csr, g, weighted, True, hub_degree=hub, respond=respond, lanes=lanes)


def test_proposal_key_widths():
    """PUSH proposals use 32-bit keys (position in the red vertex's sorted
    row, then local id) when the degrees leave room for them, 64-bit
    (weight, local id) keys otherwise; both match the oracle on lane scans,
    the round-1 head scan and the hub path.  A hub of degree 5000 in a graph of
    2^20 vertices (20-bit ids leave 12 bits of position) forces the wide keys."""
    rng = np.random.default_rng(21)
    n = 1 << 20
    ring = np.stack([np.arange(n), (np.arange(n) + 1) % n], axis=1)
    chords = rng.integers(0, n, size=(n // 2, 2))
    star = np.stack([np.zeros(5000, np.int64), 1 + rng.choice(n - 1, 5000, replace=False)], axis=1)
    e = np.concatenate([ring, chords, star])
    e = e[e[:, 0] != e[:, 1]]
    e = np.unique(np.sort(e, axis=1), axis=0)
    # ties: a few weights repeat, so the position part of a key must break them as the oracle does
    w = (rng.integers(0, 1 << 10, size=len(e)) + 1).astype(np.float32) / np.float32(1 << 10)
    wide = CSR.from_edges(n, e, w)
    narrow = mmamd.gen_rgg(1 << 16, seed=22, weighted=True)
    # 32-bit keys whose ids and positions do not fit the 4-byte head copy
    # (16 + 16 bits): the head copy keeps 8-byte {id, position} entries
    m = 1 << 16
    ring = np.stack([np.arange(m), (np.arange(m) + 1) % m], axis=1)
    star = np.stack([np.zeros(40000, np.int64), 1 + rng.choice(m - 1, 40000, replace=False)], axis=1)
    e2 = np.concatenate([ring, star])
    e2 = np.unique(np.sort(e2[e2[:, 0] != e2[:, 1]], axis=1), axis=0)
    w2 = (rng.integers(0, 1 << 8, size=len(e2)) + 1).astype(np.float32) / np.float32(1 << 8)
    mid = CSR.from_edges(m, e2, w2)
    for csr, bits in ((wide, 64), (narrow, 32), (mid, 32)):
        say("key width", bits)
        g = Graph(csr)
        info = g.info()
        assert info["push_ok"] and info["key_bits"] == bits
        for lanes, hub in ((8, 256), (4, 4096), (-1, 256)):
            check_against_oracle(csr, g, True, True, lanes=lanes, hub_degree=hub)
        g.close()


def test_weights_that_never_win_through_packed_rows():
    """Weights of -1 and below can never win a request (the reference starts
    its scan from -1.0 with a strict '>', src/matchgpu.cu:360-409); the sorted
    rows put them last with the id marked.  A graph with 32-bit keys and a packed
    head copy also scans packed 4-byte rows in the list rounds, where the mark is
    the entry's top bit: symmetric weights of -1 and -1.5 on a third of the
    edges of an RGG must give the oracle's result in every round and tile."""
    base = mmamd.gen_rgg(1 << 12, seed=31, weighted=True)
    r = base.ranges
    rows = np.repeat(np.arange(base.n), r[:, 1] - r[:, 0])
    keep = rows < base.neighbours
    e = np.stack([rows[keep], base.neighbours[keep]], axis=1).astype(np.int64)
    w = base.weights[keep].copy()
    h = (e[:, 0] * 1315423911 + e[:, 1] * 2654435761) % 9
    w[h < 2] = np.float32(-1.0)
    w[(h >= 2) & (h < 3)] = np.float32(-1.5)
    csr = CSR.from_edges(base.n, e, w)
    for keep_order in (False, True):
        g = Graph(csr, keep_order=keep_order)
        info = g.info()
        assert info["push_ok"] and info["key_bits"] == 32
        for maximal, rounds in ((True, 0), (False, 3)):
            for lanes, block, hub in ((8, 1024, 256), (2, 256, 256), (16, 64, 8)):
                say("never-win weights keep", keep_order, "maximal", maximal, "tiles", lanes, block, hub)
                exp, R, act, mat = oracle.match(csr, True, maximal, SEEDS, max_rounds=rounds, nthreads=0)
                res = g.match(kind=MM_WEIGHTED, maximal=maximal, seeds=SEEDS, max_rounds=rounds, lanes=lanes,
                              block_size=block, hub_degree=hub, respond=MM_RESPOND_PUSH)
                assert res.rounds == R
                np.testing.assert_array_equal(res.match, exp)
                np.testing.assert_array_equal(res.active[:R], act)
        g.close()


def test_packed_entries_at_their_widest():
    """Packed head and row entries hold an invalid bit, the id and the key
    position in 32 bits; a 2^16-vertex graph with a 20000-entry row needs
    16 + 15 of the 31 bits, the most the packing allows."""
    rng = np.random.default_rng(43)
    m = 1 << 16
    ring = np.stack([np.arange(m), (np.arange(m) + 1) % m], axis=1)
    star = np.stack([np.full(20000, 7, np.int64), rng.choice(np.arange(m)[np.arange(m) != 7], 20000, replace=False)],
                    axis=1)
    chords = rng.integers(0, m, size=(m, 2))
    e = np.concatenate([ring, star, chords])
    e = np.unique(np.sort(e[e[:, 0] != e[:, 1]], axis=1), axis=0)
    w = (rng.integers(0, 1 << 12, size=len(e)) + 1).astype(np.float32) / np.float32(1 << 12)
    csr = CSR.from_edges(m, e, w)
    # the packed head copy (16 bytes per vertex) comes with the packed rows (4
    # bytes per entry): dropping the head copy frees both; an unpacked head copy
    # would free only its 32 bytes per vertex
    bare = Graph(csr, heads=False)
    g = Graph(csr)
    info = g.info()
    assert info["push_ok"] and info["key_bits"] == 32
    assert info["device_bytes"] - bare.info()["device_bytes"] >= 16 * m + 4 * csr.nnz
    bare.close()
    for lanes, hub in ((8, 256), (4, 32768), (16, 64)):
        say("widest packed entries, tiles", lanes, hub)
        check_against_oracle(csr, g, True, True, lanes=lanes, hub_degree=hub)
    g.close()


def test_repeated_calls_follow_the_seeds_and_arguments():
    """Repeated calls with the same arguments reuse their option block: seeds
    changed in place between calls, and a changed argument, must still give the
    oracle's result for the values of that call."""
    csr = mmamd.gen_rgg(1 << 14, seed=61, weighted=True)
    g = Graph(csr)
    s = reference_seeds(255).copy()
    for rounds in (3, 3, 4):
        res = g.match(kind=MM_WEIGHTED, maximal=False, max_rounds=rounds, seeds=s)
        exp = oracle.match(csr, True, False, s.copy(), max_rounds=rounds, nthreads=0)[0]
        assert res.match.tolist() == exp.tolist(), rounds
        s[:] = np.roll(s, 1)  # the same array, new values
    g.close()


def test_round_one_bucket_keys_with_and_without_room():
    """Round 1 combines the proposals between rows of one bucket in LDS when
    the bucket's 32-bit keys fit beside the deferred rows, and posts every
    proposal with a global atomic when they do not: at block 1024 a graph of
    6 M vertices has 512 buckets of 11719 ids (16 KB of deferred rows + 46 KB
    of keys, over the 60 KB limit); at block 256 its 2048 buckets fit.  Both
    give the oracle's result."""
    say("round-1 bucket keys, 6 M-vertex RGG")
    csr = mmamd.gen_rgg(6_000_000, seed=31, weighted=True)
    exp, R, act, mat = oracle.match(csr, True, True, SEEDS, nthreads=0)
    g = Graph(csr)
    assert g.info()["key_bits"] == 32
    for block in (1024, 256):
        say("block", block)
        res = g.match(kind=MM_WEIGHTED, maximal=True, seeds=SEEDS, lanes=8, block_size=block)
        assert res.rounds == R
        mism = np.flatnonzero(res.match != exp)
        assert mism.size == 0, f"block {block}: {mism.size} mismatches, first at {mism[:5]}"
        np.testing.assert_array_equal(res.matched, mat)
    g.close()


def test_batch_equals_single_graph_path():
    say("batch of mixed graphs")
    rng = np.random.default_rng(3)
    graphs = []
    for k in range(40):
        n = int(2 ** rng.uniform(4, 11))
        if k % 2:
            side = max(1, int(np.sqrt(n)))
            graphs.append(mmamd.gen_grid(side, max(1, n // side), seed=k, weighted=True))
        else:
            graphs.append(mmamd.gen_rgg(n, seed=k, weighted=True))
    graphs.append(mmamd.gen_rgg(0, seed=1))
    graphs.append(mmamd.gen_grid(1, 1, seed=1))
    b = Batch(graphs)
    for maximal in (False, True):
        res, per, rounds = b.match(kind=MM_WEIGHTED, maximal=maximal, seeds=SEEDS)
        for k, csr in enumerate(graphs):
            exp, R, _, _ = oracle.match(csr, True, maximal, SEEDS)
            assert per[k].tolist() == exp.tolist(), k
            assert rounds[k] == R, (k, rounds[k], R)
        assert res.rounds == max(rounds)


def test_malformed_csr_is_rejected_on_the_device():
    """Row ranges outside the entries and neighbour ids outside the graph (or a
    batch graph's local range) are rejected before any kernel reads by them;
    the handle is not created."""
    for bad in (CSR(np.array([[0, 3], [3, 2]], np.int32), np.array([1, 0, 1], np.int32), None),   # end < begin
                CSR(np.array([[0, 1], [1, 4]], np.int32), np.array([1, 0, 1], np.int32), None),   # past the entries
                CSR(np.array([[-1, 1], [1, 2]], np.int32), np.array([1, 0], np.int32), None),     # negative begin
                CSR.from_offsets([0, 1, 2], np.array([1, 5], np.int32)),
                CSR.from_offsets([0, 1, 2], np.array([1, -1], np.int32)),
                # a row of more than 256 entries (checked by a whole block) holding id 7
                CSR.from_offsets([0, 300, 301], np.concatenate([np.ones(299), [7], [0]]).astype(np.int32))):
        with pytest.raises(mmamd.MatchError) as e:
            Graph(bad)
        assert e.value.code == mmamd.MM_ERR_INVALID and "out of bounds" in str(e.value)
    good = CSR.from_offsets([0, 1, 2], np.array([1, 0], np.int32))
    with pytest.raises(mmamd.MatchError) as e:  # graph 1's id 1 is local: past its single vertex
        Batch([good, CSR.from_offsets([0, 1], np.array([1], np.int32))])
    assert e.value.code == mmamd.MM_ERR_INVALID
    with pytest.raises(mmamd.MatchError) as e:  # graph 1's range passes its own single entry
        Batch([good, CSR(np.array([[0, 2]], np.int32), np.array([0], np.int32), None)])
    assert e.value.code == mmamd.MM_ERR_INVALID
    assert Graph(good).match(seeds=SEEDS).match.tolist() == [4, 4]


def test_empty_graph_and_bad_arguments():
    say("edge cases")
    g = Graph(CSR.from_offsets([0], np.zeros(0, np.int32)))
    res = g.match(kind=MM_WEIGHTED, maximal=True, seeds=SEEDS)
    assert res.rounds == 1 and res.match.size == 0
    res = g.match(kind=MM_WEIGHTED, maximal=False, seeds=SEEDS)
    assert res.rounds == 20
    csr = mmamd.gen_rgg(1000, seed=2)
    g = Graph(csr)
    with pytest.raises(mmamd.MatchError):
        g.match(maximal=False, seeds=SEEDS[:5])
    with pytest.raises(mmamd.MatchError):
        g.match(maximal=True, seeds=SEEDS[:2])  # cannot converge in 2 rounds
    with pytest.raises(mmamd.MatchError):
        g.match(seeds=SEEDS, lanes=3)
    with pytest.raises(mmamd.MatchError):
        g.match(seeds=SEEDS, lanes=-2)
    with pytest.raises(mmamd.MatchError):
        g.match(seeds=SEEDS, block_size=100)
    # a failed call leaves the handle usable
    exp, R, _, _ = oracle.match(csr, True, True, SEEDS)
    assert g.match(seeds=SEEDS).match.tolist() == exp.tolist()


@pytest.mark.slow
def test_rgg22_weighted_full_size_bitexact():
    """RGG n=2^22 weighted (the benchmark workload): bit-exact and valid."""
    say("rgg22 weighted full size")
    csr = mmamd.gen_rgg(1 << 22, seed=12345, weighted=True)
    g = Graph(csr)
    res = check_against_oracle(csr, g, True, True)
    ok, under, multi, invalid, matchable = oracle.test_matching(csr, res.match)
    assert ok and matchable == 0
    w_gpu, size = oracle.get_weight(csr, res.match)
    assert size == res.matched[-1]


def test_graph_replays_equal_eager_launches():
    """Captured launch sequences are replayed many times, with eager (plain
    stream) runs of other configurations in between; every replay must equal
    the eager result."""
    say("graph replay vs eager")
    csr = mmamd.gen_rgg(1 << 16, seed=17, weighted=True)
    g = Graph(csr)
    eager_max = g.match(kind=MM_WEIGHTED, maximal=True, seeds=SEEDS, graph_mode=1)
    eager_fix = g.match(kind=MM_WEIGHTED, maximal=False, max_rounds=5, seeds=SEEDS, graph_mode=1)
    for i in range(5):
        rep = g.match(kind=MM_WEIGHTED, maximal=True, seeds=SEEDS)
        assert rep.rounds == eager_max.rounds and np.array_equal(rep.match, eager_max.match), i
        np.testing.assert_array_equal(rep.matched, eager_max.matched)
        mid = g.match(kind=MM_RANDOM, maximal=True, seeds=SEEDS, graph_mode=1)
        assert oracle.test_matching(csr, mid.match)[0]
        rep = g.match(kind=MM_WEIGHTED, maximal=False, max_rounds=5, seeds=SEEDS)
        assert np.array_equal(rep.match, eager_fix.match), i


def test_consecutive_runs_alternate_counter_sets_without_leaking_state():
    """Runs switch between two counter sets; mixing eager, profiled and
    replayed runs, fixed and maximal, with and without hub rows must keep every
    result and every round count exact."""
    csr = mmamd.gen_rmat(13, seed=77, weighted=True)
    exp_max = oracle.match(csr, True, True, SEEDS, nthreads=0)
    exp_fix = oracle.match(csr, True, False, SEEDS, max_rounds=4, nthreads=0)
    g = Graph(csr)
    modes = [dict(), dict(graph_mode=1), dict(profile=True), dict(hub_degree=8), dict(hub_degree=8, graph_mode=1)]
    for k in range(12):
        kw = modes[k % len(modes)]
        maximal = k % 3 != 2
        say("run", k, kw, "maximal", maximal)
        res = g.match(kind=MM_WEIGHTED, maximal=maximal, max_rounds=0 if maximal else 4, seeds=SEEDS, **kw)
        exp, R, act, mat = exp_max if maximal else exp_fix
        assert res.rounds == R
        np.testing.assert_array_equal(res.match, exp)
        np.testing.assert_array_equal(res.active, act)
        np.testing.assert_array_equal(res.matched, mat)


def test_push_eligibility_of_crafted_csrs():
    """PUSH needs strictly ascending rows, a symmetric structure and symmetric
    weights; the check pairs every forward entry with its reverse and compares
    forward and backward counts."""
    def csr(rows, weights=None):
        offs = np.cumsum([0] + [len(r) for r in rows])
        nb = np.array([u for r in rows for u in r], np.int32)
        w = None if weights is None else np.array([x for r in weights for x in r], np.float32)
        return CSR.from_offsets(offs, nb, w)

    cases = [
        (csr([[1, 2], [0, 2], [0, 1]]), True),                                   # triangle
        (csr([[1], [2], [0]]), False),                                            # directed cycle
        (csr([[1], [], [0]]), False),                                             # counts equal, reverse missing
        (csr([[0, 1], [0, 1]]), True),                                            # self-loops ignored
        (csr([[2, 1], [0], [0]]), False),                                         # unsorted row
        (csr([[1, 2], [0], [0]], [[0.5, 0.25], [0.5], [0.75]]), False),           # asymmetric weight
        (csr([[1, 2], [0], [0]], [[0.5, 0.25], [0.5], [0.25]]), True),
    ]
    for k, (c, ok) in enumerate(cases):
        g = Graph(c, keep_order=True)
        assert bool(g.info()["push_ok"]) == ok, k
        for respond in (mmamd.MM_RESPOND_AUTO, MM_RESPOND_PULL):
            res = g.match(kind=MM_WEIGHTED, maximal=True, seeds=SEEDS, respond=respond)
            exp = oracle.match(c, True, True, SEEDS)[0]
            np.testing.assert_array_equal(res.match, exp, err_msg=str(k))


@pytest.mark.parametrize("keep", [False, True])
def test_dense_rows_exercise_the_construction_paths(keep):
    """Rows of 100-400 entries: the construction's frontier buffers overflow
    into the global list, rows fall into every length class (rank sorts of up
    to 16 and 64 entries, radix-sorted longer rows, rows past the 256-entry
    block threshold of the checks and the symmetry test)."""
    rng = np.random.default_rng(29)
    n = 1 << 17
    q = n // 4
    # ids [0, q): rows of about 8; [q, 2q): about 40; [2q, n): about 240 (a
    # search level there reaches more new vertices per block than its LDS
    # buffer holds); a few cross edges and 16 rows of about 600
    parts = [(0, q, 4 * q), (q, 2 * q, 20 * q), (2 * q, n, 240 * q)]
    a = np.concatenate([rng.integers(lo, hi, m) for lo, hi, m in parts] + [rng.integers(0, q, 4000)])
    b = np.concatenate([rng.integers(lo, hi, m) for lo, hi, m in parts] + [rng.integers(q, n, 4000)])
    hubs = rng.integers(0, n, 16)
    a = np.concatenate([a, np.repeat(hubs, 350)])
    b = np.concatenate([b, rng.integers(0, n, 16 * 350)])
    keep_e = a != b
    lo_, hi_ = np.minimum(a[keep_e], b[keep_e]), np.maximum(a[keep_e], b[keep_e])
    key = np.unique(lo_.astype(np.int64) * n + hi_)
    e = np.stack([key // n, key % n], 1)
    w = ((rng.integers(0, 1 << 24, len(e)) + 1) * 2.0 ** -24).astype(np.float32)
    csr = CSR.from_edges(n, e, w)
    deg = csr.ranges[:, 1] - csr.ranges[:, 0]
    say("dense graph", n, "nnz", csr.nnz, "max degree", int(deg.max()), "keep", keep)
    assert deg.max() > 256 and (deg <= 64).any() and ((deg > 64) & (deg <= 256)).any()
    g = Graph(csr, keep_order=keep)
    assert g.info()["push_ok"]
    for maximal in (False, True):
        res = check_against_oracle(csr, g, True, maximal)
        ok, under, multi, invalid, matchable = oracle.test_matching(csr, res.match)
        wt, size = oracle.get_weight(csr, res.match)
        chk = g.check()
        assert (chk["under"], chk["multi"], chk["invalid"], chk["matchable"], chk["size"]) == \
            (under, multi, invalid, matchable, size)
        assert chk["weight"] == wt and chk["valid"] == int(ok)
    g.close()
