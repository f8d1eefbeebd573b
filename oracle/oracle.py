"""ctypes wrapper of the CPU oracle (oracle/liborcmatch.so).  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module, and only as the checker or the timed CPU baseline — never as the
product path.  The C sources (mm_oracle.c, mm_oracle_sort.cpp) cite the
reference file:line each function restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liborcmatch.so")
_lib = None

_P = ctypes.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        h = ctypes.CDLL(LIB_PATH)
        h.orc_hash.restype = ctypes.c_uint32
        h.orc_hash.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        h.orc_colour.restype = ctypes.c_int32
        h.orc_colour.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        h.orc_hash_many.restype = None
        h.orc_hash_many.argtypes = [ctypes.c_int64, _P, _P, _P]
        h.orc_match.restype = ctypes.c_int32
        h.orc_match.argtypes = [ctypes.c_int32, _P, _P, _P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                ctypes.c_uint32, _P, ctypes.c_int32, _P, _P, _P, ctypes.c_int32]
        h.orc_get_weight.restype = None
        h.orc_get_weight.argtypes = [ctypes.c_int32, _P, _P, _P, _P, _P, _P]
        h.orc_test_matching.restype = ctypes.c_int32
        h.orc_test_matching.argtypes = [ctypes.c_int32, _P, _P, _P, _P]
        h.orc_cpu_random.restype = None
        h.orc_cpu_random.argtypes = [ctypes.c_int32, _P, _P, _P]
        h.orc_cpu_weighted.restype = None
        h.orc_cpu_weighted.argtypes = [ctypes.c_int32, _P, _P, _P, _P]
        for name in ("orc_cpu_stat_min_deg", "orc_cpu_karp_sipser", "orc_cpu_min_deg"):
            getattr(h, name).restype = None
            getattr(h, name).argtypes = [ctypes.c_int32, _P, _P, _P]
        h.orc_cpu_weighted_edge.restype = None
        h.orc_cpu_weighted_edge.argtypes = [ctypes.c_int32, ctypes.c_int64, _P, _P, _P, _P]
        h.orc_mlp_z_many.restype = None
        h.orc_mlp_z_many.argtypes = [_P, _P, ctypes.c_int64, _P]
        h.orc_blue_threshold_many.restype = None
        h.orc_blue_threshold_many.argtypes = [_P, ctypes.c_int64, _P]
        h.orc_match_learned.restype = ctypes.c_int32
        h.orc_match_learned.argtypes = [ctypes.c_int32, _P, _P, _P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.c_uint32, _P, ctypes.c_int32, _P, _P, _P, _P, ctypes.c_uint32,
                                        ctypes.c_int64, _P, _P, _P, _P, _P, _P]
        _lib = h
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _csr_arrays(csr):
    ranges = np.ascontiguousarray(csr.ranges, dtype=np.int32).reshape(-1)
    nb = np.ascontiguousarray(csr.neighbours, dtype=np.int32)
    w = None if csr.weights is None else np.ascontiguousarray(csr.weights, dtype=np.float32)
    return ranges, nb, w


def select_hash(ids, seeds) -> np.ndarray:
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    seeds = np.ascontiguousarray(np.broadcast_to(np.asarray(seeds, dtype=np.uint32), ids.shape), dtype=np.uint32)
    out = np.empty(ids.shape, np.uint32)
    lib().orc_hash_many(ids.size, _p(ids), _p(seeds), _p(out))
    return out


def colour(i: int, seed: int, barrier: int) -> int:
    return int(lib().orc_colour(i, seed, barrier))


def match(csr, weighted: bool, maximal: bool, seeds, max_rounds: int = 0, barrier: int = 0x88B81733,
          nthreads: int = 1):
    """Runs the round semantics; returns (match, rounds, active[], matched[])."""
    if max_rounds <= 0:
        max_rounds = 255 if maximal else 20
    ranges, nb, w = _csr_arrays(csr)
    if weighted and w is None:
        w = np.ones(len(nb), np.float32)
    seeds = np.ascontiguousarray(seeds, dtype=np.uint32)
    n = csr.n
    out = np.empty(n, np.int32)
    act = np.zeros(256, np.int64)
    mat = np.zeros(256, np.int64)
    R = lib().orc_match(n, _p(ranges), _p(nb), _p(w), int(weighted), int(maximal), max_rounds, barrier,
                        _p(seeds), len(seeds), _p(out), _p(act), _p(mat), nthreads)
    if R < 0:
        raise ValueError("oracle: not enough seeds")
    return out, int(R), act[:R].copy(), mat[:R].copy()


def get_weight(csr, m) -> tuple[float, int]:
    ranges, nb, w = _csr_arrays(csr)
    m = np.ascontiguousarray(m, dtype=np.int32)
    wt = ctypes.c_double()
    sz = ctypes.c_int64()
    lib().orc_get_weight(csr.n, _p(ranges), _p(nb), _p(w), _p(m), ctypes.byref(wt), ctypes.byref(sz))
    return float(wt.value), int(sz.value)


def test_matching(csr, m):
    """Returns (valid, under, multi, invalid, matchable)."""
    ranges, nb, _ = _csr_arrays(csr)
    m = np.ascontiguousarray(m, dtype=np.int32)
    counts = np.zeros(4, np.int64)
    ok = lib().orc_test_matching(csr.n, _p(ranges), _p(nb), _p(m), _p(counts))
    return bool(ok == 1), int(counts[0]), int(counts[1]), int(counts[2]), int(counts[3])


def cpu_random(csr) -> np.ndarray:
    ranges, nb, _ = _csr_arrays(csr)
    out = np.empty(csr.n, np.int32)
    lib().orc_cpu_random(csr.n, _p(ranges), _p(nb), _p(out))
    return out


def cpu_weighted(csr) -> np.ndarray:
    ranges, nb, w = _csr_arrays(csr)
    if w is None:
        w = np.ones(len(nb), np.float32)
    out = np.empty(csr.n, np.int32)
    lib().orc_cpu_weighted(csr.n, _p(ranges), _p(nb), _p(w), _p(out))
    return out


def _unweighted_host(csr, fn_name) -> np.ndarray:
    ranges, nb, _ = _csr_arrays(csr)
    out = np.empty(csr.n, np.int32)
    getattr(lib(), fn_name)(csr.n, _p(ranges), _p(nb), _p(out))
    return out


def cpu_stat_min_deg(csr) -> np.ndarray:
    """Type 1, static minimum degree (src/matchcpu.cu:428-516)."""
    return _unweighted_host(csr, "orc_cpu_stat_min_deg")


def cpu_karp_sipser(csr) -> np.ndarray:
    """Type 2, Karp-Sipser (src/matchcpu.cu:518-682)."""
    return _unweighted_host(csr, "orc_cpu_karp_sipser")


def cpu_min_deg(csr) -> np.ndarray:
    """Type 3, dynamic minimum degree (src/matchcpu.cu:304-426); quadratic, small graphs only."""
    return _unweighted_host(csr, "orc_cpu_min_deg")


def edge_list(csr):
    """Undirected edges (x, y, w) with x < y, in row order — the order the
    reference's readers leave in Graph::edges for sorted inputs."""
    ranges, nb, w = _csr_arrays(csr)
    r = ranges.reshape(-1, 2)
    src = np.repeat(np.arange(csr.n, dtype=np.int32), r[:, 1] - r[:, 0])
    keep = src < nb
    ww = np.ones(len(nb), np.float32) if w is None else w
    return src[keep].copy(), nb[keep].copy(), ww[keep].copy()


def cpu_weighted_edge(csr) -> np.ndarray:
    ex, ey, ew = edge_list(csr)
    out = np.empty(csr.n, np.int32)
    lib().orc_cpu_weighted_edge(csr.n, len(ex), _p(ex), _p(ey), _p(ew), _p(out))
    return out


# ---------------------------------------------------------------------------
# Learned selection (restates include/mm_learn.h; no reference counterpart)

NF, NH, NP = 8, 32, 1377


def mlp_z(params, x) -> np.ndarray:
    """Network outputs with the specified float32 fused multiply-add order."""
    p = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, NF)
    z = np.empty(len(x), np.float32)
    lib().orc_mlp_z_many(_p(p), _p(x), len(x), _p(z))
    return z


def blue_threshold(z) -> np.ndarray:
    z = np.ascontiguousarray(z, np.float32)
    out = np.empty(z.shape, np.uint32)
    lib().orc_blue_threshold_many(_p(z), z.size, _p(out))
    return out


def match_learned(csr, weighted: bool, maximal: bool, seeds, params=None, max_rounds: int = 0,
                  barrier: int = 0x88B81733, trace: bool = False, trace_keep: int = 65536, trace_cap: int = 0):
    """Learned rounds; returns (match, rounds, active[], matched[], samples or None)."""
    if max_rounds <= 0:
        max_rounds = 255 if maximal else 20
    ranges, nb, w = _csr_arrays(csr)
    seeds = np.ascontiguousarray(seeds, dtype=np.uint32)
    p = None if params is None else np.ascontiguousarray(params, np.float32)
    n = csr.n
    out = np.empty(n, np.int32)
    act = np.zeros(256, np.int64)
    mat = np.zeros(256, np.int64)
    cap = (trace_cap or 4 * max(n, 1)) if trace else 0
    tx = np.zeros((max(cap, 1), NF), np.float32)
    ty, tw = np.zeros(max(cap, 1), np.float32), np.zeros(max(cap, 1), np.float32)
    tv, tr = np.zeros(max(cap, 1), np.int32), np.zeros(max(cap, 1), np.int32)
    cnt = ctypes.c_int64()
    R = lib().orc_match_learned(n, _p(ranges), _p(nb), _p(w), int(weighted), int(maximal), max_rounds, barrier,
                                _p(seeds), len(seeds), _p(p), _p(out), _p(act), _p(mat), trace_keep, cap,
                                _p(tx), _p(ty), _p(tw), _p(tv), _p(tr), ctypes.byref(cnt))
    if R < 0:
        raise ValueError("oracle: not enough seeds")
    samples = None
    if trace:
        k = min(int(cnt.value), cap)
        samples = {"x": tx[:k].copy(), "y": ty[:k].copy(), "w": tw[:k].copy(), "v": tv[:k].copy(),
                   "round": tr[:k].copy(), "offered": int(cnt.value)}
    return out, int(R), act[:R].copy(), mat[:R].copy(), samples


def _unpack(params):
    p = np.asarray(params, np.float64)
    return (p[0:256].reshape(32, 8), p[256:288], p[288:1312].reshape(32, 32), p[1312:1344], p[1344:1376],
            p[1376])


def mlp_forward64(params, x):
    """float64 forward pass: (z, cache)."""
    W1, b1, W2, b2, W3, b3 = _unpack(params)
    x = np.asarray(x, np.float64).reshape(-1, NF)
    a1 = x @ W1.T + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0.0)
    z = h2 @ W3 + b3
    return z, (x, a1, h1, a2, h2)


def mlp_loss_grad64(params, x, y, w):
    """float64 loss (1/n) sum w (softplus(z) - y z) and its gradient, by
    hand-derived backpropagation (checked against finite differences)."""
    W1, b1, W2, b2, W3, b3 = _unpack(params)
    z, (x, a1, h1, a2, h2) = mlp_forward64(params, x)
    y = np.asarray(y, np.float64)
    w = np.asarray(w, np.float64)
    n = len(z)
    loss = float(np.sum(w * (np.logaddexp(0.0, z) - y * z)) / n)
    dz = w * (1.0 / (1.0 + np.exp(-z)) - y) / n
    gW3 = h2.T @ dz
    gb3 = dz.sum()
    d2 = np.outer(dz, W3) * (a2 > 0)
    gW2 = d2.T @ h1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * (a1 > 0)
    gW1 = d1.T @ x
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return loss, g
