"""Benchmark: GTEPS to >= 99% matched on a weighted random geometric graph,
with the R-MAT scale-24 graph and the 4096-graph batch measured beside it.

Headline workload (rgg22): n = 2^22 points uniform in the unit square (seed
12345), edges between points closer than r = 0.55 sqrt(ln n / n), float32
weights (k + 1) / 2^24 mirrored per undirected edge; heaviest-red-neighbour
matching with index tie-break (the reference's weighted GPU matcher, type 9),
seeds from glibc rand() after srand(12345).

A step is one matching of the graph, already resident in HBM, up to t_99: the
end of the first round whose matched-vertex count reaches 99% of the count at
convergence (the run is deterministic, so the round index is fixed).  value =
m / t_99 / 1e9 with m = undirected edges.  The tile configuration (lanes per
row, block size, hub degree) is the autotuner's pick for this step, checked to
give the same match array as the defaults.  The line also carries:
  roofline      the dominant kernel's algorithmic HBM bytes over its event-timed
                launches (on the library's stream), against 8 TB/s, with the
                bytes rocprofv3 counted for it (profiles/)
  total_ms      the reference's t0 -> t3 (bin/match.cpp:402-437): constructing
                the matcher (upload and preprocessing), one step, and copying
                the result to the host; beside it a plain host-to-device copy
                of the same CSR arrays
  cpu_baseline  the OpenMP restatement of the reference's TBB round matcher
                (type 11) on this host's cores, same graph and rounds, 32 runs
                (mean and standard deviation), with the CPU model
  cpu_matchers  the reference's sequential CPU matchers on the same graph
  learned_vs_plain  the learned selection against the plain rounds
  workloads     the same measurements for
                  rmat24     R-MAT scale 24, edge factor 16, to t_99
                  batch4096  4096 graphs (n log-uniform in [2^8, 2^16], half
                             grids, half RGGs) matched to convergence by one
                             launch sequence per round chunk

Usage: python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--workload all|rgg22|rmat24|batch4096]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "multicore-matching_amd"))

import mmamd  # noqa: E402
import workloads  # noqa: E402

HBM_PEAK_GBS = 8000.0
PROFILES = os.path.join(ROOT, "profiles")
CPU_REPS = 32  # benchmark.sh:11 runs every configuration 32 times


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_threads() -> int:
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(n, int(omp))) if omp and omp.isdigit() else n


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_info() -> dict:
    return {"cpu_model": cpu_model(), "nproc": os.cpu_count(), "threads_used": host_threads()}


def timed_reps(fn, reps: int, budget_s: float | None = None) -> dict:
    """Runs fn repeatedly (one untimed warm-up); mean and standard deviation in
    ms over `reps` runs, or over as many as fit in budget_s (at least 3)."""
    fn()
    times = []
    t_start = time.perf_counter()
    while len(times) < reps:
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
        if budget_s is not None and len(times) >= 3 and time.perf_counter() - t_start > budget_s:
            break
    ms = [t * 1e3 for t in times]
    return {"mean_ms": statistics.fmean(ms), "std_ms": statistics.stdev(ms) if len(ms) > 1 else 0.0, "reps": len(ms)}


def kernel_bytes(res: mmamd.MatchResult, rounds: int, push: bool, key_bytes: int = 8) -> dict:
    """Algorithmic HBM bytes per kernel class over the first `rounds` rounds.

    request: per blue vertex 4 (id) + 8 (range) + 4 (request write) + 8 per
             row entry (neighbour id + weight or key position); in PUSH mode
             + key_bytes (4 or 8) for the proposal's atomic on the red vertex.
    match_select: per active vertex 4 (id) + 4 (request or own key read) +
             key_bytes (partner key) + 4 (result write) + 1 (state write) + 4
             (next list entry).
    init_select: per vertex 4 (id) + 8 (range) + 1 (state) + key_bytes / 2
             (empty keys for the red half).
    respond (PULL): per red vertex 4 + 8 + 4 + 8 per row entry.
    Gathers of neighbour state are assumed to hit L2/MALL (not counted).
    round_total: the SpMV convention per active vertex, 8 (range) + 16 (match
             and request read and written) + 8 per row entry."""
    k = min(rounds, len(res.blue))
    blue, bnnz = res.blue[:k].astype(np.float64), res.blue_nnz[:k].astype(np.float64)
    act, annz = res.active[:k].astype(np.float64), res.active_nnz[:k].astype(np.float64)
    red, rnnz = act - blue, annz - bnnz
    out = {
        "request": float(np.sum(blue * (16 + (key_bytes if push else 0)) + 8 * bnnz)),
        "match_select": float(np.sum(act * (17 + key_bytes))),
        "init_select": float(res.active[0] * (13 + key_bytes / 2)) if k else 0.0,
        "respond": 0.0 if push else float(np.sum(red * 16 + 8 * rnnz)),
    }
    out["round_total"] = float(np.sum(act * (8 + 16) + 8 * annz))
    return out


def traffic_records() -> list:
    """The byte-counter records of every committed profile summary (a ```json
    block of {workload, kernel, tiles, traffic_bytes_per_launch} objects)."""
    recs = []
    for name in sorted(os.listdir(PROFILES)) if os.path.isdir(PROFILES) else []:
        if not name.endswith(".md"):
            continue
        text = open(os.path.join(PROFILES, name)).read()
        for block in text.split("```json")[1:]:
            try:
                data = json.loads(block.split("```", 1)[0])
            except ValueError:
                continue
            for r in data if isinstance(data, list) else [data]:
                if isinstance(r, dict) and "traffic_bytes_per_launch" in r:
                    recs.append(dict(r, source=f"profiles/{name}"))
    return recs


def measured_traffic(workload: str, kernel: str, tiles: dict):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 byte-counter summaries (FETCH_SIZE and WRITE_SIZE runs scaled by a
    plain-copy calibration): the record of the same tiles, else of other tiles
    (the autotuner's near-equal picks differ between runs); (None, None, None)
    when there is none."""
    recs = [r for r in traffic_records() if r.get("workload") == workload and r.get("kernel") == kernel]
    for rec in sorted(recs, key=lambda r: r.get("tiles") != tiles):
        return rec.get("traffic_bytes_per_launch"), rec.get("tiles"), rec.get("source")
    return None, None, None


def roofline(kb: dict, prof: mmamd.MatchResult, workload: str, tiles: dict) -> tuple[dict, dict]:
    kms = {k: v for k, v in prof.kernel_ms.items() if k != "counters"}
    dom = max(kms, key=kms.get)
    launches = max(1, prof.kernel_launches[dom])
    achieved = kb[dom] / (kms[dom] * 1e-3) / 1e9
    traffic, traffic_tiles, source = measured_traffic(workload, dom, tiles)
    rec = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_tiles": traffic_tiles,
           "traffic_source": source, "launches": launches, "avg_launch_us": round(kms[dom] * 1e3 / launches, 3),
           "bytes_per_launch": kb[dom] / launches}
    # every kernel class the same way: request and match_select take nearly
    # equal time on some workloads, so which one is "dominant" can change
    # between runs
    by_class = {}
    for k, ms in kms.items():
        if ms <= 0.0 or kb.get(k, 0.0) <= 0.0:
            continue
        nl = max(1, prof.kernel_launches[k])
        a = kb[k] / (ms * 1e-3) / 1e9
        by_class[k] = {"ms": round(ms, 5), "launches": nl, "avg_launch_us": round(ms * 1e3 / nl, 3),
                       "bytes_per_launch": kb[k] / nl, "achieved": round(a, 1), "frac": round(a / HBM_PEAK_GBS, 4),
                       "traffic": measured_traffic(workload, k, tiles)[0]}
    rec["by_class"] = by_class
    return rec, {k: round(v, 5) for k, v in kms.items()}


def h2d_copy_ms(csr) -> float:
    """A plain host-to-device copy of the CSR arrays (ranges, neighbours,
    weights) from pageable memory, through the HIP runtime the library uses:
    the floor of any upload (best of 3)."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    arrays = [np.ascontiguousarray(csr.ranges), np.ascontiguousarray(csr.neighbours)]
    if csr.weights is not None:
        arrays.append(np.ascontiguousarray(csr.weights))
    ptrs = []
    try:
        for a in arrays:
            p = ctypes.c_void_p()
            if hip.hipMalloc(ctypes.byref(p), a.nbytes) != 0:
                raise RuntimeError("hipMalloc failed")
            ptrs.append(p)
        best = None
        for _ in range(3):
            hip.hipDeviceSynchronize()
            t0 = time.perf_counter()
            for p, a in zip(ptrs, arrays):
                if hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) != 0:  # hipMemcpyHostToDevice
                    raise RuntimeError("hipMemcpy failed")
            hip.hipDeviceSynchronize()
            t = (time.perf_counter() - t0) * 1e3
            best = t if best is None else min(best, t)
        return best
    finally:
        for p in ptrs:
            hip.hipFree(p)


def cpu_baseline(csr, seeds, rounds: int, reps: int = CPU_REPS, budget_s: float = 30.0) -> dict:
    import oracle  # the CPU oracle is the baseline here, never the measured path

    threads = host_threads()
    t = timed_reps(lambda: oracle.match(csr, True, False, seeds, max_rounds=rounds, nthreads=threads), reps, budget_s)
    return {"value": round(csr.n_edges / (t["mean_ms"] * 1e-3) / 1e9, 4), "unit": "GTEPS", "cores": threads,
            "kind": "port", "mean_ms": round(t["mean_ms"], 3), "std_ms": round(t["std_ms"], 3), "reps": t["reps"],
            "sample": f"the same graph, first {rounds} rounds (to 99% matched): the OpenMP restatement of the "
                      f"reference's TBB weighted round matcher (type 11) on {threads} threads, {t['reps']} timed runs",
            **host_info()}


def cpu_matchers(csr, seeds, r99: int) -> dict:
    """The reference's CPU matchers on the same graph, restated in the oracle
    and timed on this host (a reported comparison, never the measured path):
    types 0 (random greedy), 4 (weighted vertex-order greedy) and 5 (edge-sorted
    greedy, src/matchcpu.cu:259-303, 695-810) single-threaded, types 1-3
    (minimum degree, Karp-Sipser, 304-682) as the library's host matchers, and
    the TBB round matcher restated with OpenMP (type 11 maximal,
    src/matchtbb.cu:567-623) on 1..all threads for the first r99 rounds plus
    one full maximal run.  32 runs each, or as many as fit in 8 s."""
    import oracle

    out = {}
    for name, fn in (("type0_random_greedy", oracle.cpu_random), ("type4_weighted_greedy", oracle.cpu_weighted),
                     ("type5_edge_sorted_greedy", oracle.cpu_weighted_edge),
                     ("type1_static_min_degree", lambda c: mmamd.host_match(c, mmamd.HOST_STAT_MIN_DEG)),
                     ("type2_karp_sipser", lambda c: mmamd.host_match(c, mmamd.HOST_KARP_SIPSER)),
                     ("type3_dynamic_min_degree", lambda c: mmamd.host_match(c, mmamd.HOST_MIN_DEG))):
        m = fn(csr)
        w, size = mmamd.get_weight(csr, m)
        t = timed_reps(lambda: fn(csr), CPU_REPS, 8.0)
        out[name] = {"matched": size, "weight": w, "mean_ms": round(t["mean_ms"], 2), "std_ms": round(t["std_ms"], 2),
                     "reps": t["reps"], "cores": 1}
    threads = host_threads()
    m, R, _, _ = oracle.match(csr, True, True, seeds, nthreads=threads)
    w, size = mmamd.get_weight(csr, m)
    t = timed_reps(lambda: oracle.match(csr, True, True, seeds, nthreads=threads), CPU_REPS, 8.0)
    out["type11_rounds_maximal"] = {"rounds": R, "matched": size, "weight": w, "mean_ms": round(t["mean_ms"], 2),
                                    "std_ms": round(t["std_ms"], 2), "reps": t["reps"], "cores": threads}
    sweep = {}
    k = 1
    while True:
        t = timed_reps(lambda: oracle.match(csr, True, False, seeds, max_rounds=r99, nthreads=k), CPU_REPS, 5.0)
        sweep[str(k)] = {"mean_ms": round(t["mean_ms"], 2), "std_ms": round(t["std_ms"], 2), "reps": t["reps"]}
        if k == threads:
            break
        k = min(k * 2, threads)
    out["type11_to_t99_by_threads"] = sweep
    out["host"] = host_info()
    return out


def tiles_of(t: dict) -> dict:
    return {"lanes": t["lanes_per_vertex"], "block_size": t["block_size"], "hub_degree": t["hub_degree"],
            "unroll": t["entries_per_thread"]}


def timed(fn, warmup: int, steps: int) -> float:
    for _ in range(warmup):
        fn()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    return (time.perf_counter() - t0) / steps


def constructed(make, reps: int = 3):
    """Builds a handle reps times (the reference's driver constructs a matcher
    for every repetition and averages, bin/match.cpp:402-437); returns the last
    handle, the median construction time and the first one (cold: the process's
    first large allocations, pageable copies and kernel code loading)."""
    times, h = [], None
    for _ in range(reps):
        if h is not None:
            h.close()
        t0 = time.perf_counter()
        h = make()
        times.append((time.perf_counter() - t0) * 1e3)
    return h, float(np.median(times)), times[0]


def run_single(args, csr, name: str, workload_desc: str, learned: bool, cpu: bool) -> dict:
    n = csr.n
    seeds = mmamd.reference_seeds(mmamd.MM_MAX_ROUNDS)
    base = dict(respond=args.respond)
    g, upload_ms, upload_first_ms = constructed(
        lambda: mmamd.Graph(csr, keep_order=args.keep_order, packed_rows=args.packed_rows, heads=args.heads))
    log(f"{name}: graph upload (incl. preprocessing) {upload_ms:.0f} ms (median of 3; first {upload_first_ms:.0f} ms)")

    full = g.match(kind=mmamd.MM_WEIGHTED, maximal=True, seeds=seeds, **base)
    final = int(full.matched[-1])
    r99 = int(np.argmax(full.matched >= 0.99 * final)) + 1
    weight, size = mmamd.get_weight(csr, full.match)
    valid = mmamd.test_matching(csr, full.match)
    log(f"{name}: maximal run rounds={full.rounds} matched={final} ({final / n:.4f}) weight={weight:.1f} "
        f"r99={r99} valid={valid} respond={full.respond} match_ms={full.match_ms:.3f}")

    if args.tune:
        tune = g.autotune(kind=mmamd.MM_WEIGHTED, maximal=False, max_rounds=r99, seeds=seeds, **base)
        if tune["mismatches"]:
            raise RuntimeError(f"autotuner candidates disagree: {tune}")
        tiles = tiles_of(tune)
    else:
        tune = None
        tiles = {"lanes": args.lanes, "block_size": args.block, "hub_degree": args.hub, "unroll": args.unroll}
    log(f"{name}: tiles {tiles} (autotune {tune})")

    def step():
        g.match(kind=mmamd.MM_WEIGHTED, maximal=False, max_rounds=r99, seeds=seeds, fetch=False, stats=False,
                **tiles, **base)

    wall = timed(step, args.warmup, args.steps)

    def step_to_host():
        g.match(kind=mmamd.MM_WEIGHTED, maximal=False, max_rounds=r99, seeds=seeds, fetch=True, stats=False,
                **tiles, **base)

    wall_host = timed(step_to_host, 1, max(3, args.steps // 4))
    ev = [g.match(kind=mmamd.MM_WEIGHTED, maximal=False, max_rounds=r99, seeds=seeds, fetch=False, **tiles,
                  **base).match_ms for _ in range(5)]
    t99_event_ms = float(np.median(ev))
    log(f"{name}: t99 wall {wall * 1e3:.4f} ms/step, device events {t99_event_ms:.4f} ms")
    # the tuned step's result is the default step's result
    a = g.match(kind=mmamd.MM_WEIGHTED, maximal=False, max_rounds=r99, seeds=seeds, **tiles, **base)
    b = g.match(kind=mmamd.MM_WEIGHTED, maximal=False, max_rounds=r99, seeds=seeds, **base)
    if not np.array_equal(a.match, b.match):
        raise RuntimeError("tuned tiles changed the match array")

    prof = g.match(kind=mmamd.MM_WEIGHTED, maximal=False, max_rounds=r99, seeds=seeds, fetch=False, profile=True,
                   **tiles, **base)
    push = prof.respond == mmamd.MM_RESPOND_PUSH
    kb = kernel_bytes(prof, r99, push, key_bytes=max(4, g.info()["key_bits"] // 8))
    default_layout = not args.keep_order and not args.packed_rows and args.heads and args.respond == 0
    roof, kms = roofline(kb, prof, f"{name}_weighted_t99" if default_layout else "", tiles)
    log(f"{name}: profile kernel_ms={kms} launches={prof.kernel_launches} bytes={kb}")
    h2d = h2d_copy_ms(csr)
    rec = {
        "metric": f"GTEPS to >=99% matched ({workload_desc}, weighted)",
        "value": round(csr.n_edges / wall / 1e9, 3),
        "unit": "GTEPS",
        "ms_per_step": round(wall * 1e3, 5),
        "config": {"workload": f"{name}: {workload_desc}" + (", r=0.55*sqrt(ln n/n)" if name.startswith("rgg") else "")
                               + f", seed {args.seed}, type 9 rounds to t_99", "n": n,
                   "m": csr.n_edges, "nnz": csr.nnz, "rounds_to_t99": r99, "rounds_to_convergence": full.rounds,
                   "respond": "push" if push else "pull", "key_bits": g.info()["key_bits"], "tiles": tiles,
                   "parallelism": "single GPU"},
        "t99_device_event_ms": round(t99_event_ms, 5),
        "gteps_device_events": round(csr.n_edges / (t99_event_ms * 1e-3) / 1e9, 3),
        "gteps_with_result_copy_to_host": round(csr.n_edges / wall_host / 1e9, 3),
        "total_ms": round(upload_ms + wall_host * 1e3, 2),
        "total_ms_parts": {"construct_upload_ms": round(upload_ms, 2), "construct_first_ms": round(upload_first_ms, 2),
                           "step_with_copy_back_ms": round(wall_host * 1e3, 4), "plain_h2d_copy_of_csr_ms": round(h2d, 2)},
        "matched_fraction": final / n,
        "matched_vertices": final,
        "matching_weight": weight,
        "valid": bool(valid[0]),
        "autotune": tune,
        "roofline": roof,
        "round_roofline_frac": round(kb["round_total"] / (t99_event_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "round_bytes": kb["round_total"],
        "kernel_ms": kms,
    }
    if learned:
        import learn

        cmp = learn.compare(csr, learn.load_params(args.params), seeds=seeds, graph=g)
        cmp["params"] = os.path.relpath(args.params, ROOT)
        rec["learned_vs_plain"] = cmp
        log(f"{name}: learned vs plain {json.dumps(cmp)}")
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        rec["cpu_baseline"] = cpu_baseline(csr, seeds, r99)
        log(f"{name}: cpu baseline {json.dumps(rec['cpu_baseline'])}")
        if name == "rgg22" and not args.no_cpu_matchers:
            rec["cpu_matchers"] = cpu_matchers(csr, seeds, r99)
            log(f"{name}: cpu matchers {json.dumps(rec['cpu_matchers'])}")
    g.close()
    return rec


def batch_cpu_baseline(graphs, seeds) -> dict:
    """The OpenMP round matcher (type 11 semantics) over every graph of the
    batch to convergence, one graph per host thread."""
    import oracle

    threads = host_threads()

    def one_pass():
        with ThreadPoolExecutor(threads) as pool:
            list(pool.map(lambda c: oracle.match(c, True, True, seeds, nthreads=1), graphs))

    t = timed_reps(one_pass, CPU_REPS, 30.0)
    m = sum(c.n_edges for c in graphs)
    return {"value": round(m / (t["mean_ms"] * 1e-3) / 1e9, 4), "unit": "GTEPS", "cores": threads, "kind": "port",
            "mean_ms": round(t["mean_ms"], 2), "std_ms": round(t["std_ms"], 2), "reps": t["reps"],
            "sample": f"all {len(graphs)} graphs to convergence: the OpenMP restatement of the reference's TBB "
                      f"round matcher (type 11), one graph per thread on {threads} threads, {t['reps']} timed passes",
            **host_info()}


def run_batch(args, cpu: bool) -> dict:
    t0 = time.perf_counter()
    specs = workloads.batch_sizes(args.batch_graphs, seed=args.seed)
    graphs = [workloads.batch_graph(*s) for s in specs]
    n_tot = sum(c.n for c in graphs)
    m_tot = sum(c.n_edges for c in graphs)
    log(f"batch: generated {len(graphs)} graphs, {n_tot} vertices, {m_tot} edges in {time.perf_counter() - t0:.1f} s")
    b, upload_ms, upload_first_ms = constructed(lambda: mmamd.Batch(graphs))
    log(f"batch: upload {upload_ms:.0f} ms (median of 3; first {upload_first_ms:.0f} ms)")
    seeds = mmamd.reference_seeds(mmamd.MM_MAX_ROUNDS)
    res, per_graph, rounds = b.match(kind=mmamd.MM_WEIGHTED, maximal=True, seeds=seeds)
    log(f"batch: maximal rounds={res.rounds} (per graph {rounds.min()}..{rounds.max()}) match_ms={res.match_ms:.3f}")
    # every graph's result is the single-graph path's (a seeded sample of graphs, with the smallest and largest)
    rng = np.random.default_rng(args.seed)
    sizes = np.array([c.n for c in graphs])
    pick = sorted(set(rng.choice(len(graphs), size=min(args.batch_check, len(graphs)), replace=False).tolist()
                      + [int(np.argmin(sizes)), int(np.argmax(sizes))]))
    bad = 0
    for k in pick:
        single = mmamd.Graph(graphs[k]).match(kind=mmamd.MM_WEIGHTED, maximal=True, seeds=seeds)
        bad += int(not np.array_equal(single.match, per_graph[k]) or single.rounds != rounds[k])
    check_ms = []
    for _ in range(3):
        t0 = time.perf_counter()
        chk, gw, gs, gv = b.check()  # getWeight / testMatching of every graph, on the device
        check_ms.append((time.perf_counter() - t0) * 1e3)
    log(f"batch: {len(pick)} graphs against the single-graph path: {bad} differ; device check {chk} "
        f"({int(gv.sum())} of {len(gv)} graphs valid) in {min(check_ms):.1f} ms")
    tune = b.autotune(kind=mmamd.MM_WEIGHTED, maximal=True, seeds=seeds) if args.tune else None
    tiles = tiles_of(tune) if tune else {"lanes": args.lanes, "block_size": args.block, "hub_degree": args.hub,
                                         "unroll": args.unroll}
    log(f"batch: tiles {tiles} (autotune {tune})")

    def step():
        b.match(kind=mmamd.MM_WEIGHTED, maximal=True, seeds=seeds, fetch=False, **tiles)

    wall = timed(step, args.warmup, args.steps)
    ev = float(np.median([b.match(kind=mmamd.MM_WEIGHTED, maximal=True, seeds=seeds, fetch=False,
                                  **tiles)[0].match_ms for _ in range(5)]))

    def step_to_host():
        b.match(kind=mmamd.MM_WEIGHTED, maximal=True, seeds=seeds, fetch=True, **tiles)

    wall_host = timed(step_to_host, 1, 3)
    tuned, _, _ = b.match(kind=mmamd.MM_WEIGHTED, maximal=True, seeds=seeds, **tiles)
    if not np.array_equal(tuned.match, res.match):
        raise RuntimeError("tuned tiles changed the batch result")
    prof, _, _ = b.match(kind=mmamd.MM_WEIGHTED, maximal=True, seeds=seeds, fetch=False, profile=True, **tiles)
    kb = kernel_bytes(prof, prof.rounds, prof.respond == mmamd.MM_RESPOND_PUSH,
                      key_bytes=max(4, b.info()["key_bits"] // 8))
    roof, kms = roofline(kb, prof, "batch4096_weighted_maximal", tiles)
    log(f"batch: profile kernel_ms={kms} launches={prof.kernel_launches} bytes={kb}")
    rec = {
        "metric": "GTEPS, 4096-graph batch matched to convergence (weighted)",
        "value": round(m_tot / wall / 1e9, 3), "unit": "GTEPS", "ms_per_step": round(wall * 1e3, 4),
        "config": {"workload": f"batch{len(graphs)}: n log-uniform in [2^8, 2^16], even graphs 2-D grids "
                               f"(4-neighbour), odd graphs RGG r=0.55*sqrt(ln n/n), seed {args.seed}; type 9 "
                               f"rounds to convergence", "graphs": len(graphs), "n": n_tot, "m": m_tot,
                   "rounds": res.rounds, "rounds_per_graph": [int(rounds.min()), int(np.median(rounds)),
                                                              int(rounds.max())],
                   "tiles": tiles, "parallelism": "single GPU"},
        "device_event_ms": round(ev, 4), "gteps_device_events": round(m_tot / (ev * 1e-3) / 1e9, 3),
        "total_ms": round(upload_ms + wall_host * 1e3, 2),
        "total_ms_parts": {"construct_upload_ms": round(upload_ms, 2), "construct_first_ms": round(upload_first_ms, 2),
                           "step_with_copy_back_ms": round(wall_host * 1e3, 4)},
        "roofline": roof,
        "round_roofline_frac": round(kb["round_total"] / (ev * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "round_bytes": kb["round_total"],
        "kernel_ms": kms,
        "matched_vertices": chk["size"], "matching_weight": chk["weight"], "valid": bool(chk["valid"]),
        "graphs_valid": int(gv.sum()), "still_matchable": chk["matchable"], "device_check_ms": round(min(check_ms), 2),
        "device_check_first_ms": round(check_ms[0], 2),
        "graphs_checked_against_single_path": len(pick), "graphs_differing": bad, "autotune": tune,
    }
    b.close()
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        rec["cpu_baseline"] = batch_cpu_baseline(graphs, seeds)
        log(f"batch: cpu baseline {json.dumps(rec['cpu_baseline'])}")
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=("all", "rgg22", "rmat24", "batch4096"), default="all",
                    help="all: the headline rgg22 line with rmat24 and batch4096 records inside it")
    ap.add_argument("--log2n", type=int, default=22, help="rgg workload size")
    ap.add_argument("--scale", type=int, default=24, help="rmat workload scale")
    ap.add_argument("--batch-graphs", type=int, default=4096)
    ap.add_argument("--batch-check", type=int, default=32, help="batch graphs checked against the single path")
    ap.add_argument("--seed", type=int, default=12345)
    ap.add_argument("--no-tune", dest="tune", action="store_false", help="fixed tiles instead of the autotuner")
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--hub", type=int, default=0)
    ap.add_argument("--unroll", type=int, default=0, help="row entries per thread of the tile kernels (4, 8, 16)")
    ap.add_argument("--respond", type=int, default=mmamd.MM_RESPOND_AUTO)
    ap.add_argument("--no-cpu", action="store_true", help="skip every CPU timing")
    ap.add_argument("--no-cpu-matchers", action="store_true", help="skip the sequential CPU matchers")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-selection comparison")
    ap.add_argument("--params", default=os.path.join(ROOT, "multicore-matching_amd", "mlp_params.json"),
                    help="network parameters of the learned selection (learn.py train)")
    ap.add_argument("--keep-order", action="store_true", help="no internal locality renumbering")
    ap.add_argument("--packed-rows", action="store_true", help="device rows back to back (no line alignment)")
    ap.add_argument("--no-heads", dest="heads", action="store_false",
                    help="no dense copy of the rows' first entries (round 1 scans the rows)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="run only a plain device copy of this many MiB (calibrates rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch

        x = torch.ones(args.copy_probe << 18, dtype=torch.int32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": args.copy_probe << 20, "launches": 3}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: the matcher runs on one GPU; sharding a graph over several GPUs is not implemented. "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if mmamd.lib().mm_device_count() < 1:
        print("bench.py: no GPU visible to the HIP runtime", file=sys.stderr)
        sys.exit(1)

    # HIP runtime and code-object initialisation happen before any timed
    # construction (the reference's t0 also follows its context set-up)
    mmamd.Graph(mmamd.gen_rgg(4096, seed=1, weighted=True)).close()
    cpu = not args.no_cpu
    t0 = time.perf_counter()
    if args.workload == "batch4096":
        line = run_batch(args, cpu)
    elif args.workload == "rmat24":
        csr = workloads.rmat(args.scale, seed=args.seed)
        log(f"generated R-MAT scale {args.scale}: n={csr.n} nnz={csr.nnz} in {time.perf_counter() - t0:.2f} s")
        line = run_single(args, csr, f"rmat{args.scale}", f"R-MAT scale {args.scale}, edge factor 16, "
                                                          f"(0.57, 0.19, 0.19, 0.05)", False, cpu)
    else:
        csr = workloads.rgg(args.log2n, seed=args.seed)
        log(f"generated RGG n={csr.n} nnz={csr.nnz} m={csr.n_edges} in {time.perf_counter() - t0:.2f} s")
        line = run_single(args, csr, f"rgg{args.log2n}", f"2^{args.log2n} RGG", not args.no_learned, cpu)
        del csr
        if args.workload == "all":
            others = {}
            t1 = time.perf_counter()
            rm = workloads.rmat(args.scale, seed=args.seed)
            log(f"generated R-MAT scale {args.scale}: n={rm.n} nnz={rm.nnz} in {time.perf_counter() - t1:.2f} s")
            others[f"rmat{args.scale}"] = run_single(args, rm, f"rmat{args.scale}",
                                                     f"R-MAT scale {args.scale}, edge factor 16, "
                                                     f"(0.57, 0.19, 0.19, 0.05)", False, cpu)
            del rm
            others[f"batch{args.batch_graphs}"] = run_batch(args, cpu)
            line["workloads"] = others
    line.update({"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
                 "scaling": "strong", "vs_baseline": None, "dtype": "int32+f32cmp",
                 "data": "synthetic (seeded, generated on the host)"})
    line["config"]["parallelism"] = "single GPU"
    line["bench_s"] = round(time.perf_counter() - t0, 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
