"""ctypes binding of libmmamd.so, the MI355X greedy matcher's C ABI (include/mm_match.h,
include/mm_learn.h).

Python is only the test/bench host here: it hands numpy arrays to the C ABI and
reads results back.  The matcher itself runs in the HIP kernels of
csrc/mm_device.hip; there is no CPU fallback — loading fails loudly when the
library has not been built.

Mirrors the reference's operator surface (BasFaggingerAuer/Multicore-Matching):
  Graph(...)             <- Graph CSR + GraphMatchingGPU ctor (src/graph.h:41-60, src/matchgpu.cu:39-141)
  Graph.match(...)       <- GraphMatchingGPU*::performMatching (src/matchgpu.cu:453-851)
  TYPE_* codes           <- bin/match.cpp:145-148 (types 6-9)
  reference_seeds(...)   <- srand(12345) then one rand() per round (bin/match.cpp:382, src/matchgpu.cu:509)
  Graph.match_learned    <- no counterpart: network-chosen blue probabilities (include/mm_learn.h)
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmmamd.so")

MM_OK, MM_ERR_INVALID, MM_ERR_NOMEM, MM_ERR_HIP, MM_ERR_SEEDS = 0, 1, 2, 3, 4
MM_RANDOM, MM_WEIGHTED = 0, 1
MM_RESPOND_AUTO, MM_RESPOND_PULL, MM_RESPOND_PUSH = 0, 1, 2
MM_NR_MATCH_ROUNDS = 20
MM_NR_MAX_MATCH_ROUNDS = 255
MM_MAX_ROUNDS = 255
MM_DEFAULT_BARRIER = 0x88B81733
MM_CHECK_NOALLOC_MAX_N = 4194304
MM_CREATE_KEEP_ORDER = 1
MM_CREATE_PACKED_ROWS = 2
MM_CREATE_NO_HEADS = 4
MM_LANES_TILES = -1
MM_KERNEL_CLASSES = 5
KERNEL_CLASS_NAMES = ("init_select", "request", "respond", "match_select", "counters")

# Reference matcher type codes of the GPU path (bin/match.cpp:145-148).
TYPE_GPU_RANDOM, TYPE_GPU_RANDOM_MAXIMAL, TYPE_GPU_WEIGHTED, TYPE_GPU_WEIGHTED_MAXIMAL = 6, 7, 8, 9
TYPE_TO_KIND = {6: (MM_RANDOM, 0), 7: (MM_RANDOM, 1), 8: (MM_WEIGHTED, 0), 9: (MM_WEIGHTED, 1)}


class MatchError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[mm error {code}] {msg}")
        self.code = code


class _Opts(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("maximal", ctypes.c_int32),
        ("max_rounds", ctypes.c_int32),
        ("barrier", ctypes.c_uint32),
        ("seeds", ctypes.c_void_p),  # const uint32_t *
        ("n_seeds", ctypes.c_int32),
        ("block_size", ctypes.c_int32),
        ("lanes_per_vertex", ctypes.c_int32),
        ("hub_degree", ctypes.c_int32),
        ("entries_per_thread", ctypes.c_int32),
        ("respond", ctypes.c_int32),
        ("profile", ctypes.c_int32),
        ("graph_mode", ctypes.c_int32),
    ]


_R1 = MM_MAX_ROUNDS + 1


class _Stats(ctypes.Structure):
    _fields_ = [
        ("rounds", ctypes.c_int32),
        ("respond_used", ctypes.c_int32),
        ("lanes_used", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("match_ms", ctypes.c_float),
        ("kernel_ms", ctypes.c_float * MM_KERNEL_CLASSES),
        ("kernel_launches", ctypes.c_int32 * MM_KERNEL_CLASSES),
        ("active", ctypes.c_int64 * _R1),
        ("blue", ctypes.c_int64 * _R1),
        ("matched", ctypes.c_int64 * _R1),
        ("active_nnz", ctypes.c_int64 * _R1),
        ("blue_nnz", ctypes.c_int64 * _R1),
    ]


class _Info(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32),
        ("weighted", ctypes.c_int32),
        ("nnz", ctypes.c_int64),
        ("max_degree", ctypes.c_int32),
        ("push_ok", ctypes.c_int32),
        ("device_bytes", ctypes.c_int64),
        ("key_bits", ctypes.c_int32),
        ("iso_settled", ctypes.c_int32),
    ]


class _Csr(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32),
        ("nnz", ctypes.c_int64),
        ("ranges", ctypes.POINTER(ctypes.c_int32)),
        ("neighbours", ctypes.POINTER(ctypes.c_int32)),
        ("weights", ctypes.POINTER(ctypes.c_float)),
    ]


MM_MLP_FEATURES, MM_MLP_HIDDEN, MM_MLP_PARAMS = 8, 32, 1377
MLP_OFF = {"W1": 0, "b1": 256, "W2": 288, "b2": 1312, "W3": 1344, "b3": 1376}
MM_LEARN_DRAW_TWEAK, MM_LEARN_TRACE_TWEAK = 0x5BD1E995, 0x2545F491


class _LearnOpts(ctypes.Structure):
    _fields_ = [
        ("params", ctypes.c_void_p),
        ("trace", ctypes.c_int32),
        ("trace_keep", ctypes.c_uint32),
        ("trace_cap", ctypes.c_int64),
        ("trace_x", ctypes.c_void_p),
        ("trace_y", ctypes.c_void_p),
        ("trace_w", ctypes.c_void_p),
        ("trace_v", ctypes.c_void_p),
        ("trace_round", ctypes.c_void_p),
        ("trace_count", ctypes.c_int64),
        ("trace_offered", ctypes.c_int64),
    ]


class _TrainOpts(ctypes.Structure):
    _fields_ = [
        ("epochs", ctypes.c_int32),
        ("batch", ctypes.c_int32),
        ("lr", ctypes.c_float),
        ("beta1", ctypes.c_float),
        ("beta2", ctypes.c_float),
        ("eps", ctypes.c_float),
        ("seed", ctypes.c_uint64),
    ]


class _TuneOpts(ctypes.Structure):
    _fields_ = [("reps", ctypes.c_int32), ("verify", ctypes.c_int32)]


class _TuneResult(ctypes.Structure):
    _fields_ = [
        ("block_size", ctypes.c_int32),
        ("lanes_per_vertex", ctypes.c_int32),
        ("hub_degree", ctypes.c_int32),
        ("entries_per_thread", ctypes.c_int32),
        ("best_ms", ctypes.c_float),
        ("default_ms", ctypes.c_float),
        ("candidates", ctypes.c_int32),
        ("mismatches", ctypes.c_int32),
    ]


def _tune_dict(r: _TuneResult) -> dict:
    return {k: getattr(r, k) for k, _ in _TuneResult._fields_}


class _Check(ctypes.Structure):
    _fields_ = [("under", ctypes.c_int64), ("multi", ctypes.c_int64), ("invalid", ctypes.c_int64),
                ("matchable", ctypes.c_int64), ("size", ctypes.c_int64), ("weight", ctypes.c_double),
                ("valid", ctypes.c_int32)]


def _check_dict(c: _Check) -> dict:
    return {k: getattr(c, k) for k, _ in _Check._fields_}


# Every symbol include/*.h declares, with its ctypes signature.
_P = ctypes.c_void_p
_I32, _I64, _U64, _D = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double
SYMBOLS = {
    "mm_last_error": (ctypes.c_char_p, []),
    "mm_device_count": (_I32, []),
    "mm_graph_create": (_I32, [_I32, _I32, _I64, _P, _P, _P, ctypes.POINTER(_P)]),
    "mm_graph_create_ex": (_I32, [_I32, _I32, _I64, _P, _P, _P, ctypes.c_uint32, ctypes.POINTER(_P)]),
    "mm_graph_destroy": (_I32, [_P]),
    "mm_graph_get_info": (_I32, [_P, ctypes.POINTER(_Info)]),
    "mm_match": (_I32, [_P, ctypes.POINTER(_Opts), _P, ctypes.POINTER(_Stats)]),
    "mm_graph_copy_match": (_I32, [_P, _P]),
    "mm_batch_create": (_I32, [_I32, _I32, _P, _P, _P, _P, _P, ctypes.POINTER(_P)]),
    "mm_batch_destroy": (_I32, [_P]),
    "mm_batch_get_info": (_I32, [_P, ctypes.POINTER(_Info)]),
    "mm_batch_match": (_I32, [_P, ctypes.POINTER(_Opts), _P, _P, ctypes.POINTER(_Stats)]),
    "mm_gen_rgg": (_I32, [_I32, _D, _U64, _I32, ctypes.POINTER(_Csr)]),
    "mm_gen_rmat": (_I32, [_I32, _I32, _D, _D, _D, _U64, _I32, ctypes.POINTER(_Csr)]),
    "mm_gen_grid": (_I32, [_I32, _I32, _U64, _I32, ctypes.POINTER(_Csr)]),
    "mm_csr_free": (None, [ctypes.POINTER(_Csr)]),
    "mm_get_weight": (_I32, [_I32, _P, _P, _P, _P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_I64)]),
    "mm_test_matching": (_I32, [_I32, _P, _P, _P, _P]),
    "mm_host_match": (_I32, [_I32, _I32, _I64, _P, _P, _P, _P]),
    "mm_graph_check": (_I32, [_P, ctypes.POINTER(_Check)]),
    "mm_graph_set_match": (_I32, [_P, _P]),
    "mm_graph_check_noalloc": (_I32, [_P, ctypes.POINTER(_Check)]),
    "mm_batch_check": (_I32, [_P, ctypes.POINTER(_Check), _P, _P, _P]),
    "mm_batch_set_match": (_I32, [_P, _P]),
    "mm_autotune": (_I32, [_P, ctypes.POINTER(_Opts), ctypes.POINTER(_TuneOpts), ctypes.POINTER(_TuneResult)]),
    "mm_batch_autotune": (_I32, [_P, ctypes.POINTER(_Opts), ctypes.POINTER(_TuneOpts),
                                 ctypes.POINTER(_TuneResult)]),
    "mm_match_learned": (_I32, [_P, ctypes.POINTER(_Opts), ctypes.POINTER(_LearnOpts), _P, ctypes.POINTER(_Stats)]),
    "mm_mlp_init": (_I32, [_U64, _P]),
    "mm_mlp_forward": (_I32, [_I32, _P, _P, _I64, _P]),
    "mm_mlp_loss_grad": (_I32, [_I32, _P, _P, _P, _P, _I64, ctypes.POINTER(_D), _P]),
    "mm_mlp_train": (_I32, [_I32, _P, _P, _P, _I64, ctypes.POINTER(_TrainOpts), _P, _P]),
}

_lib = None


def lib():
    """Loads libmmamd.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          "(make -C multicore-matching_amd)")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SYMBOLS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def _check(rc: int):
    if rc != MM_OK:
        raise MatchError(rc, lib().mm_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def reference_seeds(count: int, srand_seed: int = 12345) -> np.ndarray:
    """One glibc rand() per round after srand(srand_seed), as the reference
    driver draws them (bin/match.cpp:382, src/matchgpu.cu:509)."""
    libc = ctypes.CDLL("libc.so.6")
    libc.rand.restype = ctypes.c_int
    libc.srand(ctypes.c_uint(srand_seed))
    return np.array([libc.rand() for _ in range(count)], dtype=np.uint32)


@dataclass
class CSR:
    """Host CSR in the reference's Graph layout (src/graph.h:41-60)."""
    ranges: np.ndarray          # int32[n, 2]: begin, end
    neighbours: np.ndarray      # int32[nnz]
    weights: np.ndarray | None  # float32[nnz] or None

    @property
    def n(self) -> int:
        return int(self.ranges.shape[0])

    @property
    def nnz(self) -> int:
        return int(self.neighbours.shape[0])

    @property
    def n_edges(self) -> int:
        return self.nnz // 2

    @staticmethod
    def from_offsets(offsets, neighbours, weights=None) -> "CSR":
        offsets = np.asarray(offsets, dtype=np.int64)
        ranges = np.stack([offsets[:-1], offsets[1:]], axis=1).astype(np.int32)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        return CSR(np.ascontiguousarray(ranges), np.ascontiguousarray(neighbours, dtype=np.int32), w)

    @staticmethod
    def from_edges(n: int, edges, weights=None, sort_rows: bool = True) -> "CSR":
        """Undirected edge list -> symmetric CSR (both directions stored)."""
        edges = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
        w = np.ones(len(edges), np.float32) if weights is None else np.asarray(weights, np.float32)
        src = np.concatenate([edges[:, 0], edges[:, 1]])
        dst = np.concatenate([edges[:, 1], edges[:, 0]])
        ww = np.concatenate([w, w])
        order = np.lexsort((dst, src)) if sort_rows else np.argsort(src, kind="stable")
        src, dst, ww = src[order], dst[order], ww[order]
        counts = np.bincount(src, minlength=n)
        offsets = np.concatenate([[0], np.cumsum(counts)])
        return CSR.from_offsets(offsets, dst, None if weights is None else ww)


def _csr_from_struct(c: _Csr) -> CSR:
    n, nnz = int(c.n), int(c.nnz)
    ranges = np.ctypeslib.as_array(c.ranges, shape=(max(n, 1) * 2,))[: 2 * n].copy().reshape(n, 2)
    nb = np.ctypeslib.as_array(c.neighbours, shape=(max(nnz, 1),))[:nnz].copy()
    w = None
    if c.weights:
        w = np.ctypeslib.as_array(c.weights, shape=(max(nnz, 1),))[:nnz].copy()
    lib().mm_csr_free(ctypes.byref(c))
    return CSR(ranges, nb, w)


def get_weight(csr: CSR, match: np.ndarray) -> tuple[float, int]:
    """(weight, matched vertices) of a match array — GraphMatching::getWeight."""
    r = np.ascontiguousarray(csr.ranges, np.int32)
    nb = np.ascontiguousarray(csr.neighbours, np.int32)
    w = None if csr.weights is None else np.ascontiguousarray(csr.weights, np.float32)
    m = np.ascontiguousarray(match, np.int32)
    wt, sz = ctypes.c_double(), ctypes.c_int64()
    _check(lib().mm_get_weight(csr.n, _ptr(r), _ptr(nb), _ptr(w), _ptr(m), ctypes.byref(wt), ctypes.byref(sz)))
    return float(wt.value), int(sz.value)


# Sequential host matcher types (bin/match.cpp:136-143).
HOST_RANDOM, HOST_STAT_MIN_DEG, HOST_KARP_SIPSER, HOST_MIN_DEG, HOST_WEIGHTED, HOST_WEIGHTED_EDGE = range(6)


def host_match(csr: CSR, type_code: int) -> np.ndarray:
    """One of the reference's sequential host matchers (types 0-5) run by the
    library on the host; returns the match array in the reference encoding."""
    r = np.ascontiguousarray(csr.ranges, np.int32)
    nb = np.ascontiguousarray(csr.neighbours, np.int32)
    w = None if csr.weights is None else np.ascontiguousarray(csr.weights, np.float32)
    out = np.empty(csr.n, np.int32)
    _check(lib().mm_host_match(type_code, csr.n, len(nb), _ptr(r), _ptr(nb), _ptr(w), _ptr(out)))
    return out


def test_matching(csr: CSR, match: np.ndarray):
    """GraphMatching::testMatching: (valid, under, multi, invalid, still matchable)."""
    r = np.ascontiguousarray(csr.ranges, np.int32)
    nb = np.ascontiguousarray(csr.neighbours, np.int32)
    m = np.ascontiguousarray(match, np.int32)
    c = np.zeros(4, np.int64)
    rc = lib().mm_test_matching(csr.n, _ptr(r), _ptr(nb), _ptr(m), _ptr(c))
    return rc == MM_OK, int(c[0]), int(c[1]), int(c[2]), int(c[3])


test_matching.__test__ = False  # not a pytest test


def gen_rgg(n: int, seed: int = 12345, weighted: bool = True, radius_scale: float = 0.55) -> CSR:
    c = _Csr()
    _check(lib().mm_gen_rgg(n, radius_scale, seed, int(weighted), ctypes.byref(c)))
    return _csr_from_struct(c)


def gen_rmat(scale: int, edge_factor: int = 16, a=0.57, b=0.19, c=0.19, seed: int = 12345,
             weighted: bool = True) -> CSR:
    s = _Csr()
    _check(lib().mm_gen_rmat(scale, edge_factor, a, b, c, seed, int(weighted), ctypes.byref(s)))
    return _csr_from_struct(s)


def gen_grid(rows: int, cols: int, seed: int = 12345, weighted: bool = True) -> CSR:
    c = _Csr()
    _check(lib().mm_gen_grid(rows, cols, seed, int(weighted), ctypes.byref(c)))
    return _csr_from_struct(c)


@dataclass
class MatchResult:
    match: np.ndarray | None
    rounds: int
    match_ms: float
    respond: int
    lanes: int
    active: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    blue: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    matched: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    active_nnz: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    blue_nnz: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    kernel_ms: dict = field(default_factory=dict)
    kernel_launches: dict = field(default_factory=dict)


def _address(a: np.ndarray) -> int:
    """Data address of a contiguous array (the buffer route is several times
    faster than numpy's ctypes attribute, which matters at one call per
    sub-millisecond matching; read-only arrays take the attribute)."""
    try:
        return ctypes.addressof(ctypes.c_char.from_buffer(a))
    except (TypeError, ValueError):
        return a.ctypes.data


def _make_opts(kind, maximal, seeds, max_rounds, barrier, block_size, lanes, hub_degree, respond, profile,
               graph_mode=0, unroll=0):
    seeds = np.ascontiguousarray(seeds, dtype=np.uint32)
    o = _Opts(kind, int(maximal), max_rounds, barrier, _address(seeds) if len(seeds) else None, len(seeds),
              block_size, lanes, hub_degree, int(unroll), respond, int(profile), int(graph_mode))
    return o, seeds


def _result(st: _Stats, match) -> MatchResult:
    R = int(st.rounds)
    k = min(R, MM_MAX_ROUNDS)
    return MatchResult(
        match=match, rounds=R, match_ms=float(st.match_ms), respond=int(st.respond_used),
        lanes=int(st.lanes_used),
        active=np.array(st.active[:k], np.int64), blue=np.array(st.blue[:k], np.int64),
        matched=np.array(st.matched[:k], np.int64), active_nnz=np.array(st.active_nnz[:k], np.int64),
        blue_nnz=np.array(st.blue_nnz[:k], np.int64),
        kernel_ms={KERNEL_CLASS_NAMES[i]: float(st.kernel_ms[i]) for i in range(MM_KERNEL_CLASSES)},
        kernel_launches={KERNEL_CLASS_NAMES[i]: int(st.kernel_launches[i]) for i in range(MM_KERNEL_CLASSES)},
    )


class Graph:
    """Device-resident CSR graph (mm_graph)."""

    def __init__(self, csr: CSR, device: int = 0, keep_order: bool = False, packed_rows: bool = False,
                 heads: bool = True):
        self.csr = csr
        self.n, self.nnz = csr.n, csr.nnz
        h = ctypes.c_void_p()
        ranges = np.ascontiguousarray(csr.ranges, dtype=np.int32)
        nb = np.ascontiguousarray(csr.neighbours, dtype=np.int32)
        w = None if csr.weights is None else np.ascontiguousarray(csr.weights, dtype=np.float32)
        flags = ((MM_CREATE_KEEP_ORDER if keep_order else 0) | (MM_CREATE_PACKED_ROWS if packed_rows else 0)
                 | (0 if heads else MM_CREATE_NO_HEADS))
        _check(lib().mm_graph_create_ex(device, self.n, self.nnz, _ptr(ranges), _ptr(nb), _ptr(w), flags,
                                        ctypes.byref(h)))
        self._h = h

    def info(self) -> dict:
        i = _Info()
        _check(lib().mm_graph_get_info(self._h, ctypes.byref(i)))
        return {k: getattr(i, k) for k, _ in _Info._fields_}

    def match(self, kind: int = MM_WEIGHTED, maximal: bool = True, seeds=None, max_rounds: int = 0,
              barrier: int = MM_DEFAULT_BARRIER, block_size: int = 0, lanes: int = 0, hub_degree: int = 0,
              respond: int = MM_RESPOND_AUTO, profile: bool = False, fetch: bool = True,
              stats: bool = True, graph_mode: int = 0, unroll: int = 0) -> MatchResult | None:
        """Runs one matching.  fetch=False leaves the match array on the device;
        stats=False skips the counter read-back and returns None."""
        if seeds is None:
            seeds = reference_seeds(MM_MAX_ROUNDS)
        # A repeated call with the same arguments and the same uint32 seeds
        # array reuses its option block (the library reads the seeds' current
        # values through it; the reference held here keeps numpy from resizing
        # the array under it).
        key = (kind, maximal, max_rounds, barrier, block_size, lanes, hub_degree, respond, profile, graph_mode,
               unroll)
        c = getattr(self, "_opts_cache", None)
        if c is not None and c[0] == key and c[2] is seeds:
            o = c[1]
        else:
            o, keep = _make_opts(kind, maximal, seeds, max_rounds, barrier, block_size, lanes, hub_degree, respond,
                                 profile, graph_mode, unroll)
            self._opts_cache = (key, o, keep) if keep is seeds else None
        st = _Stats() if stats else None
        out = np.empty(self.n, np.int32) if fetch else None
        _check(lib().mm_match(self._h, ctypes.byref(o), _ptr(out), ctypes.byref(st) if stats else None))
        return _result(st, out) if stats else None

    def match_learned(self, params=None, kind: int = MM_WEIGHTED, maximal: bool = True, seeds=None,
                      max_rounds: int = 0, barrier: int = MM_DEFAULT_BARRIER, block_size: int = 0,
                      fetch: bool = True, trace: bool = False, trace_keep: int = 65536, trace_cap: int | None = None):
        """Rounds with network-chosen blue probabilities (mm_match_learned);
        params=None runs the plain colouring through the same kernels.  With
        trace=True returns (result, samples) where samples is a dict of x, y, w,
        v, round arrays ordered by (round, v) and 'offered'."""
        if seeds is None:
            seeds = reference_seeds(MM_MAX_ROUNDS)
        o, keep = _make_opts(kind, maximal, seeds, max_rounds, barrier, block_size, 0, 0, 0, False)
        lo = _LearnOpts()
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        if p is not None and p.shape != (MM_MLP_PARAMS,):
            raise ValueError(f"params must hold {MM_MLP_PARAMS} floats")
        lo.params = None if p is None else p.ctypes.data
        bufs = None
        if trace:
            cap = self.n * 4 if trace_cap is None else int(trace_cap)
            bufs = {"x": np.zeros((cap, MM_MLP_FEATURES), np.float32), "y": np.zeros(cap, np.float32),
                    "w": np.zeros(cap, np.float32), "v": np.zeros(cap, np.int32), "round": np.zeros(cap, np.int32)}
            lo.trace, lo.trace_keep, lo.trace_cap = 1, trace_keep, cap
            lo.trace_x, lo.trace_y, lo.trace_w = bufs["x"].ctypes.data, bufs["y"].ctypes.data, bufs["w"].ctypes.data
            lo.trace_v, lo.trace_round = bufs["v"].ctypes.data, bufs["round"].ctypes.data
        st = _Stats()
        out = np.empty(self.n, np.int32) if fetch else None
        _check(lib().mm_match_learned(self._h, ctypes.byref(o), ctypes.byref(lo), _ptr(out), ctypes.byref(st)))
        del keep
        res = _result(st, out)
        if not trace:
            return res
        k = int(lo.trace_count)
        samples = {key: a[:k] for key, a in bufs.items()}
        samples["offered"] = int(lo.trace_offered)
        return res, samples

    def autotune(self, kind: int = MM_WEIGHTED, maximal: bool = True, seeds=None, max_rounds: int = 0,
                 barrier: int = MM_DEFAULT_BARRIER, respond: int = MM_RESPOND_AUTO, reps: int = 3,
                 verify: bool = True) -> dict:
        """Fastest block size / lanes per row / hub degree for this graph and run (mm_autotune)."""
        if seeds is None:
            seeds = reference_seeds(MM_MAX_ROUNDS)
        o, keep = _make_opts(kind, maximal, seeds, max_rounds, barrier, 0, 0, 0, respond, False)
        r = _TuneResult()
        _check(lib().mm_autotune(self._h, ctypes.byref(o), ctypes.byref(_TuneOpts(reps, int(verify))),
                                 ctypes.byref(r)))
        del keep
        return _tune_dict(r)

    def check(self) -> dict:
        """getWeight / testMatching of the current device result, on the device (mm_graph_check)."""
        c = _Check()
        _check(lib().mm_graph_check(self._h, ctypes.byref(c)))
        return _check_dict(c)

    def check_noalloc(self) -> dict:
        """As check(), with the library's static scratch instead of allocated memory
        (mm_graph_check_noalloc; graphs of up to MM_CHECK_NOALLOC_MAX_N vertices)."""
        c = _Check()
        _check(lib().mm_graph_check_noalloc(self._h, ctypes.byref(c)))
        return _check_dict(c)

    def set_match(self, match):
        """Replaces the device result with a host array in the caller's order (mm_graph_set_match)."""
        m = np.ascontiguousarray(match, np.int32)
        if m.shape != (self.n,):
            raise ValueError("match must hold one value per vertex")
        _check(lib().mm_graph_set_match(self._h, _ptr(m)))

    def match_type(self, type_code: int, seeds=None, **kw) -> MatchResult:
        kind, maximal = TYPE_TO_KIND[type_code]
        return self.match(kind=kind, maximal=bool(maximal), seeds=seeds, **kw)

    def copy_match(self) -> np.ndarray:
        out = np.empty(self.n, np.int32)
        _check(lib().mm_graph_copy_match(self._h, _ptr(out)))
        return out

    def close(self):
        if getattr(self, "_h", None):
            lib().mm_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many independent graphs matched by one launch sequence (mm_batch)."""

    def __init__(self, graphs: list[CSR], device: int = 0):
        self.graphs = graphs
        vo = np.zeros(len(graphs) + 1, np.int32)
        eo = np.zeros(len(graphs) + 1, np.int64)
        for k, g in enumerate(graphs):
            vo[k + 1] = vo[k] + g.n
            eo[k + 1] = eo[k] + g.nnz
        self.vertex_offsets, self.edge_offsets = vo, eo
        ranges = np.ascontiguousarray(np.concatenate([g.ranges.reshape(-1) for g in graphs]) if graphs
                                      else np.zeros(0, np.int32), dtype=np.int32)
        nb = np.ascontiguousarray(np.concatenate([g.neighbours for g in graphs]) if graphs
                                  else np.zeros(0, np.int32), dtype=np.int32)
        weighted = bool(graphs) and all(g.weights is not None for g in graphs)
        if not weighted and any(g.weights is not None for g in graphs):
            raise ValueError("Batch: either every graph has weights or none has (a weighted matching would "
                             "otherwise run on unit weights for the graphs that have them)")
        w = np.ascontiguousarray(np.concatenate([g.weights for g in graphs]), dtype=np.float32) if weighted else None
        h = ctypes.c_void_p()
        _check(lib().mm_batch_create(device, len(graphs), _ptr(vo), _ptr(eo), _ptr(ranges), _ptr(nb), _ptr(w),
                                     ctypes.byref(h)))
        self._h = h
        self.total_n = int(vo[-1])

    def match(self, kind: int = MM_WEIGHTED, maximal: bool = True, seeds=None, max_rounds: int = 0,
              barrier: int = MM_DEFAULT_BARRIER, block_size: int = 0, lanes: int = 0, hub_degree: int = 0,
              respond: int = MM_RESPOND_AUTO, profile: bool = False, fetch: bool = True, graph_mode: int = 0,
              unroll: int = 0):
        if seeds is None:
            seeds = reference_seeds(MM_MAX_ROUNDS)
        o, keep = _make_opts(kind, maximal, seeds, max_rounds, barrier, block_size, lanes, hub_degree, respond,
                             profile, graph_mode, unroll)
        st = _Stats()
        out = np.empty(self.total_n, np.int32) if fetch else None
        rounds = np.zeros(len(self.graphs), np.int32)
        _check(lib().mm_batch_match(self._h, ctypes.byref(o), _ptr(out), _ptr(rounds), ctypes.byref(st)))
        del keep
        res = _result(st, out)
        per_graph = None
        if out is not None:
            per_graph = [out[self.vertex_offsets[k]:self.vertex_offsets[k + 1]] for k in range(len(self.graphs))]
        return res, per_graph, rounds

    def autotune(self, kind: int = MM_WEIGHTED, maximal: bool = True, seeds=None, max_rounds: int = 0,
                 barrier: int = MM_DEFAULT_BARRIER, respond: int = MM_RESPOND_AUTO, reps: int = 3,
                 verify: bool = True) -> dict:
        if seeds is None:
            seeds = reference_seeds(MM_MAX_ROUNDS)
        o, keep = _make_opts(kind, maximal, seeds, max_rounds, barrier, 0, 0, 0, respond, False)
        r = _TuneResult()
        _check(lib().mm_batch_autotune(self._h, ctypes.byref(o), ctypes.byref(_TuneOpts(reps, int(verify))),
                                       ctypes.byref(r)))
        del keep
        return _tune_dict(r)

    def info(self) -> dict:
        i = _Info()
        _check(lib().mm_batch_get_info(self._h, ctypes.byref(i)))
        return {k: getattr(i, k) for k, _ in _Info._fields_}

    def _set_match(self, match):
        """Replaces the device result with the graphs' arrays concatenated (mm_batch_set_match)."""
        m = np.ascontiguousarray(match, np.int32)
        if m.shape != (self.total_n,):
            raise ValueError("match must hold one value per vertex of the batch")
        _check(lib().mm_batch_set_match(self._h, _ptr(m)))

    def check(self):
        """(totals, per-graph weight, per-graph matched vertices, per-graph valid) on the device."""
        c = _Check()
        k = len(self.graphs)
        gw, gs, gv = np.zeros(k, np.float64), np.zeros(k, np.int64), np.zeros(k, np.int32)
        _check(lib().mm_batch_check(self._h, ctypes.byref(c), _ptr(gw), _ptr(gs), _ptr(gv)))
        return _check_dict(c), gw, gs, gv.astype(bool)

    def close(self):
        if getattr(self, "_h", None):
            lib().mm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mlp_init(seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases (mm_mlp_init)."""
    p = np.zeros(MM_MLP_PARAMS, np.float32)
    _check(lib().mm_mlp_init(seed, _ptr(p)))
    return p


def mlp_forward(params, x, device: int = 0) -> np.ndarray:
    """Network outputs z of feature rows x[n, 8] on the device (mm_mlp_forward)."""
    p = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, MM_MLP_FEATURES)
    z = np.zeros(len(x), np.float32)
    _check(lib().mm_mlp_forward(device, _ptr(p), _ptr(x), len(x), _ptr(z)))
    return z


def mlp_loss_grad(params, x, y, w, device: int = 0):
    """(loss, gradient) of (1/n) sum w (softplus(z) - y z) by the device kernels."""
    p = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, MM_MLP_FEATURES)
    y = np.ascontiguousarray(y, np.float32)
    w = np.ascontiguousarray(w, np.float32)
    g = np.zeros(MM_MLP_PARAMS, np.float32)
    loss = ctypes.c_double()
    _check(lib().mm_mlp_loss_grad(device, _ptr(p), _ptr(x), _ptr(y), _ptr(w), len(x), ctypes.byref(loss), _ptr(g)))
    return float(loss.value), g


def mlp_train(x, y, w, params, epochs: int = 4, batch: int = 4096, lr: float = 3e-3, seed: int = 1,
              device: int = 0):
    """Minibatch Adam on the device (mm_mlp_train): (trained params, per-epoch mean loss)."""
    p = np.array(params, np.float32, copy=True)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, MM_MLP_FEATURES)
    y = np.ascontiguousarray(y, np.float32)
    w = np.ascontiguousarray(w, np.float32)
    t = _TrainOpts(epochs, batch, lr, 0.9, 0.999, 1e-8, seed)
    el = np.zeros(epochs, np.float32)
    _check(lib().mm_mlp_train(device, _ptr(x), _ptr(y), _ptr(w), len(x), ctypes.byref(t), _ptr(p), _ptr(el)))
    return p, el
