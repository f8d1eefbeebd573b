"""GPU checks of the learned selection (include/mm_learn.h) through the C ABI:
the learned rounds against the CPU oracle bit for bit (match array, rounds,
per-round counts, traces), the network kernels against the oracle and a
float64 statement, training determinism, and what the trained network gains
over the plain rounds on a graph it was not trained on."""
import sys

import numpy as np
import pytest

import learn
import mmamd
import oracle
from mmamd import MM_RANDOM, MM_WEIGHTED, Graph, reference_seeds

pytestmark = pytest.mark.gpu

SEEDS = reference_seeds(255)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def shuffled_rows(csr, seed):
    """Rows in random order: the CSR no longer has ascending rows, so the
    rounds take the exact respond scan (PULL) instead of proposal atomics."""
    rng = np.random.default_rng(seed)
    nb, w = csr.neighbours.copy(), csr.weights.copy()
    for i in range(csr.n):
        a, b = csr.ranges[i]
        p = rng.permutation(b - a)
        nb[a:b], w[a:b] = nb[a:b][p], w[a:b][p]
    return mmamd.CSR(csr.ranges.copy(), nb, w)


def graphs():
    return [("rgg13", mmamd.gen_rgg(1 << 13, seed=31, weighted=True)),
            ("rmat12", mmamd.gen_rmat(12, seed=32, weighted=True)),
            ("grid", mmamd.gen_grid(70, 50, seed=33, weighted=True)),
            ("rgg12u", mmamd.gen_rgg(1 << 12, seed=34, weighted=False)),
            ("rgg12_shuffled", shuffled_rows(mmamd.gen_rgg(1 << 12, seed=35, weighted=True), 36))]


@pytest.mark.parametrize("keep", [False, True])
def test_plain_colouring_through_learned_kernels_equals_mm_match(keep):
    for name, csr in graphs():
        g = Graph(csr, keep_order=keep)
        for kind in (MM_RANDOM, MM_WEIGHTED):
            for maximal in (False, True):
                say(name, "keep", keep, "kind", kind, "maximal", maximal)
                a = g.match(kind=kind, maximal=maximal, seeds=SEEDS)
                b = g.match_learned(None, kind=kind, maximal=maximal, seeds=SEEDS)
                assert a.rounds == b.rounds
                np.testing.assert_array_equal(a.match, b.match)
                np.testing.assert_array_equal(a.active, b.active)
                np.testing.assert_array_equal(a.matched, b.matched)
        g.close()


@pytest.mark.parametrize("keep", [False, True])
def test_learned_rounds_and_traces_equal_the_oracle(keep):
    for name, csr in graphs():
        g = Graph(csr, keep_order=keep)
        for ps in (0, 1):
            p = mmamd.mlp_init(100 + ps)
            for kind in (MM_RANDOM, MM_WEIGHTED):
                say(name, "keep", keep, "params", ps, "kind", kind)
                res, s = g.match_learned(p, kind=kind, maximal=True, seeds=SEEDS, trace=True, trace_keep=40000)
                exp, R, act, mat, es = oracle.match_learned(csr, kind == MM_WEIGHTED, True, SEEDS, params=p,
                                                            trace=True, trace_keep=40000)
                assert res.rounds == R
                np.testing.assert_array_equal(res.match, exp)
                np.testing.assert_array_equal(res.active, act)
                np.testing.assert_array_equal(res.matched, mat)
                assert s["offered"] == es["offered"]
                for key in ("x", "y", "w", "v", "round"):
                    np.testing.assert_array_equal(s[key], es[key], err_msg=key)
                assert mmamd.test_matching(csr, res.match)[0]
            # fixed round counts stop mid-way with the colours of the last select
            res = g.match_learned(p, kind=MM_WEIGHTED, maximal=False, seeds=SEEDS, max_rounds=3)
            exp = oracle.match_learned(csr, True, False, SEEDS, params=p, max_rounds=3)[0]
            np.testing.assert_array_equal(res.match, exp)
        g.close()


def test_plain_traces_equal_the_oracle():
    csr = mmamd.gen_rgg(1 << 14, seed=41, weighted=True)
    g = Graph(csr)
    res, s = g.match_learned(None, kind=MM_WEIGHTED, maximal=True, seeds=SEEDS, trace=True)
    _, R, _, mat, es = oracle.match_learned(csr, True, True, SEEDS, trace=True)
    assert s["offered"] == es["offered"] == mat[-1]
    for key in ("x", "y", "w", "v", "round"):
        np.testing.assert_array_equal(s[key], es[key], err_msg=key)


def test_forward_kernel_equals_oracle_bitwise():
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 2, (20000, 8)).astype(np.float32)
    for seed in range(3):
        p = mmamd.mlp_init(seed)
        z = mmamd.mlp_forward(p, x)
        np.testing.assert_array_equal(z, oracle.mlp_z(p, x))
        z64, _ = oracle.mlp_forward64(p, x)
        np.testing.assert_allclose(z, z64, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("n", [1, 63, 64, 65, 5000, 100000])
def test_gradient_kernels_match_float64_oracle(n):
    rng = np.random.default_rng(n)
    x = rng.uniform(0, 1.5, (n, 8)).astype(np.float32)
    y = (rng.uniform(size=n) < 0.5).astype(np.float32)
    w = rng.uniform(0.05, 1.0, n).astype(np.float32)
    p = mmamd.mlp_init(9)
    loss, g = mmamd.mlp_loss_grad(p, x, y, w)
    loss64, g64 = oracle.mlp_loss_grad64(p, x, y, w)
    say("n", n, "loss", loss, loss64, "max |g|", np.abs(g64).max())
    assert abs(loss - loss64) <= 1e-5 * max(1.0, abs(loss64))
    tol = 1e-4 * np.abs(g64).max() + 1e-7
    np.testing.assert_allclose(g, g64, rtol=1e-3, atol=tol)


def test_training_is_reproducible_and_lowers_the_loss():
    csr = mmamd.gen_rgg(1 << 14, seed=51, weighted=True)
    g = Graph(csr)
    _, s = g.match_learned(None, kind=MM_WEIGHTED, maximal=True, seeds=SEEDS, trace=True)
    p0 = mmamd.mlp_init(1)
    l0, _ = mmamd.mlp_loss_grad(p0, s["x"], s["y"], s["w"])
    p1, losses1 = mmamd.mlp_train(s["x"], s["y"], s["w"], p0, epochs=5, batch=512, lr=3e-3, seed=3)
    p2, losses2 = mmamd.mlp_train(s["x"], s["y"], s["w"], p0, epochs=5, batch=512, lr=3e-3, seed=3)
    np.testing.assert_array_equal(p1, p2)
    np.testing.assert_array_equal(losses1, losses2)
    l1, _ = mmamd.mlp_loss_grad(p1, s["x"], s["y"], s["w"])
    say("loss before", l0, "after", l1, "per epoch", losses1)
    assert l1 < l0 and losses1[-1] < losses1[0]


def test_trained_network_beats_the_plain_rounds_on_an_unseen_graph():
    """Ablation: the same rounds with and without the network.  Trained on
    plain traces of 16 small RGG / R-MAT graphs, evaluated on a larger RGG."""
    x, y, w = learn.collect_traces(learn.training_graphs(16, 12, 14, seed=77), samples_per_graph=8192,
                                   seeds=SEEDS)
    p, losses = learn.train(x, y, w, epochs=6, batch=1024)
    say("samples", len(x), "loss per epoch", losses)
    csr = mmamd.gen_rgg(1 << 16, seed=12345, weighted=True)
    cmp = learn.compare(csr, p, seeds=SEEDS)
    say(cmp)
    untrained = learn.compare(csr, mmamd.mlp_init(1), seeds=SEEDS)
    say("untrained", untrained["learned"])
    assert cmp["plain"]["valid"] and cmp["learned"]["valid"]
    assert cmp["learned"]["rounds"] < cmp["plain"]["rounds"]
    assert cmp["learned"]["weight"] > cmp["plain"]["weight"]
    assert cmp["learned"]["rounds_to_plain_weight"] is not None
    assert cmp["learned"]["rounds_to_plain_weight"] < cmp["plain"]["rounds"]
    assert cmp["learned"]["weight"] > untrained["learned"]["weight"]
