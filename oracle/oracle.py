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
        h.orc_cpu_weighted_edge.restype = None
        h.orc_cpu_weighted_edge.argtypes = [ctypes.c_int32, ctypes.c_int64, _P, _P, _P, _P]
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
