"""CPU checks of the C ABI library: it loads, exports every symbol that
include/mm_match.h declares, rejects malformed input on the host, and its
host-side generators produce the documented graphs.  No GPU calls."""
import os
import re

import numpy as np
import pytest

import mmamd
from mmamd import CSR

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(mm_[a-z_0-9]+)\s*\(", text, re.M))
    return names


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert {"mm_graph_create", "mm_match", "mm_batch_match", "mm_gen_rgg", "mm_last_error"} <= names
    lib = mmamd.lib()
    for name in names:
        assert hasattr(lib, name), name
    assert names == set(mmamd.SYMBOLS), names ^ set(mmamd.SYMBOLS)


def test_no_device_here_is_reported_not_faked():
    if mmamd.lib().mm_device_count() > 0:
        pytest.skip("a GPU is visible")
    csr = mmamd.gen_rgg(100, seed=1)
    with pytest.raises(mmamd.MatchError) as e:
        mmamd.Graph(csr)
    assert e.value.code == mmamd.MM_ERR_HIP


def coords(n, seed):
    """The RGG point coordinates, restated with numpy (SplitMix64 finaliser)."""
    m64 = (1 << 64) - 1

    def mix(x):
        x = (x + 0x9E3779B97F4A7C15) & m64
        x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & m64
        x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & m64
        return x ^ (x >> 31)

    pseed = mix(seed ^ 0x504F494E54530000)
    ms = mix(pseed)
    xs = np.array([(mix(ms ^ (2 * i)) >> 11) * 2.0 ** -53 for i in range(n)])
    ys = np.array([(mix(ms ^ (2 * i + 1)) >> 11) * 2.0 ** -53 for i in range(n)])
    return xs, ys


def test_rgg_matches_brute_force_distance_graph():
    n = 700
    csr = mmamd.gen_rgg(n, seed=4242, weighted=True)
    x, y = coords(n, 4242)
    r = 0.55 * np.sqrt(np.log(n) / n)
    d2 = (x[:, None] - x[None, :]) ** 2 + (y[:, None] - y[None, :]) ** 2
    adj = (d2 < r * r) & ~np.eye(n, dtype=bool)
    for i in range(n):
        a, b = csr.ranges[i]
        assert csr.neighbours[a:b].tolist() == np.flatnonzero(adj[i]).tolist()


def check_symmetric_sorted(csr):
    rows = [csr.neighbours[a:b] for a, b in csr.ranges]
    wts = [csr.weights[a:b] for a, b in csr.ranges] if csr.weights is not None else None
    for i, row in enumerate(rows):
        assert np.all(np.diff(row) > 0)
        assert i not in row
        for j, u in enumerate(row):
            k = np.searchsorted(rows[u], i)
            assert rows[u][k] == i
            if wts is not None:
                assert wts[u][k] == wts[i][j]
                assert 0.0 < wts[i][j] <= 1.0 and (wts[i][j] * 2 ** 24) == int(wts[i][j] * 2 ** 24)


def test_generators_symmetric_sorted_mirrored_and_thread_independent(monkeypatch):
    for csr in (mmamd.gen_rgg(3000, seed=1), mmamd.gen_rmat(10, 8, seed=2), mmamd.gen_grid(13, 17, seed=3)):
        check_symmetric_sorted(csr)
    grid = mmamd.gen_grid(3, 4, seed=3, weighted=False)
    assert grid.neighbours[grid.ranges[5][0]:grid.ranges[5][1]].tolist() == [1, 4, 6, 9]
    assert grid.nnz == 2 * (3 * 3 + 2 * 4)
    monkeypatch.setenv("MM_GEN_THREADS", "1")
    a = mmamd.gen_rmat(12, 16, seed=9)
    b = mmamd.gen_rgg(20000, seed=9)
    monkeypatch.setenv("MM_GEN_THREADS", "7")
    c = mmamd.gen_rmat(12, 16, seed=9)
    d = mmamd.gen_rgg(20000, seed=9)
    for p, q in ((a, c), (b, d)):
        assert np.array_equal(p.ranges, q.ranges) and np.array_equal(p.neighbours, q.neighbours)
        assert np.array_equal(p.weights, q.weights)


def test_rgg_mean_degree_matches_the_workload():
    csr = mmamd.gen_rgg(1 << 16, seed=12345)
    assert 10.0 < csr.nnz / csr.n < 11.0  # pi r^2 n = 0.3025 pi ln n, minus boundary loss


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys

    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2 and "one GPU" in p.stderr and p.stdout == ""
