"""mm_graph_check_noalloc (the result check on the library's static scratch,
no device allocation) against mm_graph_check (handle-owned scratch) through
the C ABI: the same counts and weight, bit for bit, on two calls in a row --
the second on a smaller graph, so stale contents of the shared scratch from
the first call would show -- for valid and planted-invalid arrays, both vertex
orders; and the size limit."""
import sys

import numpy as np
import pytest

import mmamd
from check_util import plant_invalid
from mmamd import MM_WEIGHTED, Graph, reference_seeds

pytestmark = pytest.mark.gpu

SEEDS = reference_seeds(255)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("keep", [False, True])
def test_noalloc_check_equals_allocating_check_on_two_calls(keep):
    rng = np.random.default_rng(31)
    big = mmamd.gen_rgg(1 << 20, seed=41, weighted=True)
    small = mmamd.gen_rgg(1 << 13, seed=42, weighted=True)
    for name, csr in (("rgg20", big), ("rgg13", small)):
        g = Graph(csr, keep_order=keep)
        res = g.match(kind=MM_WEIGHTED, maximal=True, seeds=SEEDS)
        for what, arr in (("result", res.match), ("planted", plant_invalid(res.match, csr, rng))):
            g.set_match(arr)
            say(name, "keep", keep, what)
            a = g.check_noalloc()
            b = g.check()
            assert a == b, (name, what, a, b)
            if what == "planted":
                assert not a["valid"] and a["under"]
        g.close()


def test_noalloc_check_full_size_and_limit():
    csr = mmamd.gen_rgg(mmamd.MM_CHECK_NOALLOC_MAX_N, seed=12345, weighted=True)
    g = Graph(csr)
    g.match(kind=MM_WEIGHTED, maximal=False, max_rounds=5, seeds=SEEDS)
    say("rgg22 at the limit")
    assert g.check_noalloc() == g.check()
    g.close()
    over = mmamd.gen_grid(2049, 2048, seed=3, weighted=True)
    assert over.n > mmamd.MM_CHECK_NOALLOC_MAX_N
    g = Graph(over)
    g.match(kind=MM_WEIGHTED, maximal=True, seeds=SEEDS)
    with pytest.raises(mmamd.MatchError):
        g.check_noalloc()
    assert g.check()["valid"]
    g.close()
