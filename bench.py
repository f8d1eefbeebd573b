"""Benchmark: GTEPS to >= 99% matched on a weighted random geometric graph.

Default workload (rgg22): n = 2^22 points uniform in the unit square (seed
12345), edges between points closer than r = 0.55 sqrt(ln n / n), float32
weights (k + 1) / 2^24 mirrored per undirected edge; heaviest-red-neighbour
matching with index tie-break (the reference's weighted GPU matcher, type 9),
seeds from glibc rand() after srand(12345).

A step is one matching of the graph, already resident in HBM, up to t_99: the
end of the first round whose matched-vertex count reaches 99% of the count at
convergence (the run is deterministic, so the round index is fixed).  value =
m / t_99 / 1e9 with m = undirected edges.  The tile configuration (lanes per
row, block size, hub degree) is the autotuner's pick for this step, checked to
give the same match array as the defaults.  The JSON line also carries the
roofline of the dominant kernel (HIP events on the library's stream), the CPU
baseline (the OpenMP restatement of the reference's TBB round matcher, type
11, timed on this host), and the learned selection against the plain rounds on
the same graph: rounds, matched vertices, weight and device time to
convergence, and the rounds and time the learned run needs to reach the plain
run's final weight.

Other workloads (one JSON line each, not the headline):
  --workload rmat24    R-MAT scale 24, edge factor 16, the same step and metric
  --workload batch4096 4096 graphs (n log-uniform in [2^8, 2^16], half grids,
                       half RGGs) matched to convergence by one launch
                       sequence per round chunk; GTEPS over all their edges

Usage: python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--workload rgg22]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "multicore-matching_amd"))

import mmamd  # noqa: E402
import workloads  # noqa: E402

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "rgg22_weighted_t99_kernels.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


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
    Gathers of neighbour state are assumed to hit L2/MALL (not counted)."""
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


def measured_traffic(workload: str, kernel: str, tiles: dict):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 byte-counter
    summary (FETCH_SIZE and WRITE_SIZE runs scaled by a plain-copy calibration)
    and the tiles they were measured with: the record of the same tiles, else
    the record of the same workload and kernel under other tiles (the
    autotuner's near-equal picks differ from run to run), else (None, None)."""
    try:
        text = open(TRAFFIC_FILE).read()
        recs = json.loads(text.split("```json", 1)[1].split("```", 1)[0])
    except (OSError, IndexError, ValueError):
        return None, None
    recs = [r for r in (recs if isinstance(recs, list) else [recs])
            if r.get("workload") == workload and r.get("kernel") == kernel]
    for rec in sorted(recs, key=lambda r: r.get("tiles") != tiles):
        return rec.get("traffic_bytes_per_launch"), rec.get("tiles")
    return None, None


def cpu_baseline(csr, seeds, r99: int, t_budget_s: float = 20.0) -> dict:
    import oracle  # the CPU oracle is the baseline here, never the measured path

    threads = os.cpu_count() or 1
    # warm-up pass also sizes the repetitions to the budget
    t0 = time.perf_counter()
    oracle.match(csr, True, False, seeds, max_rounds=r99, nthreads=threads)
    one = time.perf_counter() - t0
    reps = max(1, min(16, int(t_budget_s / max(one, 1e-3))))
    t0 = time.perf_counter()
    for _ in range(reps):
        oracle.match(csr, True, False, seeds, max_rounds=r99, nthreads=threads)
    t = (time.perf_counter() - t0) / reps
    return {"value": csr.n_edges / t / 1e9, "unit": "GTEPS", "cores": threads, "kind": "port",
            "sample": f"same graph, first {r99} rounds (to 99% matched), OpenMP restatement of the reference's "
                      f"TBB weighted round matcher (type 11), {reps} timed runs, {t * 1e3:.1f} ms each"}


def cpu_matchers(csr, seeds, r99: int) -> dict:
    """The reference's CPU matchers on the same graph, restated in the oracle
    and timed on this host (a reported comparison, never the measured path):
    types 0 (random greedy), 4 (weighted vertex-order greedy) and 5 (edge-sorted
    greedy, src/matchcpu.cu:259-303, 695-810) single-threaded, types 1-3
    (minimum degree, Karp-Sipser, 304-682) as the library's host matchers, and the TBB
    round matcher restated with OpenMP (type 11 maximal, src/matchtbb.cu:567-623)
    on 1..all cores for the first r99 rounds plus one full maximal run."""
    import oracle

    out = {}
    for name, fn in (("type0_random_greedy", oracle.cpu_random), ("type4_weighted_greedy", oracle.cpu_weighted),
                     ("type5_edge_sorted_greedy", oracle.cpu_weighted_edge)):
        t0 = time.perf_counter()
        m = fn(csr)
        dt = time.perf_counter() - t0
        w, size = mmamd.get_weight(csr, m)
        out[name] = {"matched": size, "weight": w, "ms": round(dt * 1e3, 2), "cores": 1}
    # types 1-3 (minimum degree, Karp-Sipser): the library's host matchers
    for name, t in (("type1_static_min_degree", mmamd.HOST_STAT_MIN_DEG), ("type2_karp_sipser", mmamd.HOST_KARP_SIPSER),
                    ("type3_dynamic_min_degree", mmamd.HOST_MIN_DEG)):
        t0 = time.perf_counter()
        m = mmamd.host_match(csr, t)
        dt = time.perf_counter() - t0
        w, size = mmamd.get_weight(csr, m)
        out[name] = {"matched": size, "weight": w, "ms": round(dt * 1e3, 2), "cores": 1}
    threads = os.cpu_count() or 1
    t0 = time.perf_counter()
    m, R, _, mat = oracle.match(csr, True, True, seeds, nthreads=threads)
    dt = time.perf_counter() - t0
    w, size = mmamd.get_weight(csr, m)
    out["type11_rounds_maximal"] = {"rounds": R, "matched": size, "weight": w, "ms": round(dt * 1e3, 2),
                                    "cores": threads}
    sweep = {}
    k = 1
    while k <= threads:
        t0 = time.perf_counter()
        oracle.match(csr, True, False, seeds, max_rounds=r99, nthreads=k)
        sweep[str(k)] = round((time.perf_counter() - t0) * 1e3, 2)
        k = k * 2 if k * 2 <= threads or k == threads else threads
    out["type11_to_t99_ms_by_threads"] = sweep
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


def run_single(args, csr, name: str, workload_desc: str) -> dict:
    n = csr.n
    t0 = time.perf_counter()
    g = mmamd.Graph(csr, keep_order=args.keep_order, packed_rows=args.packed_rows, heads=args.heads)
    log(f"graph upload (incl. renumbering) {time.perf_counter() - t0:.2f} s")
    seeds = mmamd.reference_seeds(mmamd.MM_MAX_ROUNDS)
    base = dict(respond=args.respond)

    full = g.match(kind=mmamd.MM_WEIGHTED, maximal=True, seeds=seeds, **base)
    final = int(full.matched[-1])
    r99 = int(np.argmax(full.matched >= 0.99 * final)) + 1
    weight, size = mmamd.get_weight(csr, full.match)
    valid = mmamd.test_matching(csr, full.match)
    log(f"maximal run: rounds={full.rounds} matched={final} ({final / n:.4f}) weight={weight:.1f} "
        f"r99={r99} valid={valid} respond={full.respond} match_ms={full.match_ms:.3f}")

    if args.tune:
        tune = g.autotune(kind=mmamd.MM_WEIGHTED, maximal=False, max_rounds=r99, seeds=seeds, **base)
        if tune["mismatches"]:
            raise RuntimeError(f"autotuner candidates disagree: {tune}")
        tiles = tiles_of(tune)
    else:
        tune = None
        tiles = {"lanes": args.lanes, "block_size": args.block, "hub_degree": args.hub, "unroll": args.unroll}
    log(f"tiles {tiles} (autotune {tune})")

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
    log(f"t99: wall {wall * 1e3:.4f} ms/step, device events {t99_event_ms:.4f} ms")
    # the tuned step's result is the default step's result
    a = g.match(kind=mmamd.MM_WEIGHTED, maximal=False, max_rounds=r99, seeds=seeds, **tiles, **base)
    b = g.match(kind=mmamd.MM_WEIGHTED, maximal=False, max_rounds=r99, seeds=seeds, **base)
    if not np.array_equal(a.match, b.match):
        raise RuntimeError("tuned tiles changed the match array")

    prof = g.match(kind=mmamd.MM_WEIGHTED, maximal=False, max_rounds=r99, seeds=seeds, fetch=False, profile=True,
                   **tiles, **base)
    push = prof.respond == mmamd.MM_RESPOND_PUSH
    kb = kernel_bytes(prof, r99, push, key_bytes=max(4, g.info()["key_bits"] // 8))
    kms = {k: v for k, v in prof.kernel_ms.items() if k != "counters"}
    dom = max(kms, key=kms.get)
    launches = prof.kernel_launches[dom]
    achieved = kb[dom] / (kms[dom] * 1e-3) / 1e9
    log(f"profile: kernel_ms={kms} launches={prof.kernel_launches} bytes={kb}")
    traffic, traffic_tiles = None, None
    if not args.keep_order and not args.packed_rows and args.heads and args.respond == 0:
        traffic, traffic_tiles = measured_traffic(f"{name}_weighted_t99", dom, tiles)

    line = {
        "metric": f"GTEPS to >=99% matched ({workload_desc}, weighted)",
        "value": round(csr.n_edges / wall / 1e9, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 5),
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "int32+f32cmp",
        "data": "synthetic (seeded, generated on the host)",
        "config": {"workload": f"{name}: {workload_desc}" + (", r=0.55*sqrt(ln n/n)" if name.startswith("rgg") else "")
                               + f", seed {args.seed}, type 9 rounds to t_99", "n": n,
                   "m": csr.n_edges, "nnz": csr.nnz, "rounds_to_t99": r99, "rounds_to_convergence": full.rounds,
                   "respond": "push" if push else "pull", "tiles": tiles, "parallelism": "single GPU"},
        "t99_device_event_ms": round(t99_event_ms, 5),
        "gteps_device_events": round(csr.n_edges / (t99_event_ms * 1e-3) / 1e9, 3),
        "gteps_with_result_copy_to_host": round(csr.n_edges / wall_host / 1e9, 3),
        "matched_fraction": final / n,
        "matched_vertices": final,
        "matching_weight": weight,
        "valid": bool(valid[0]),
        "autotune": tune,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_tiles": traffic_tiles, "launches": launches, "avg_launch_us": round(kms[dom] * 1e3 / max(1, launches), 3),
                     "bytes_per_launch": kb[dom] / max(1, launches)},
        "round_roofline_frac": round(kb["round_total"] / (t99_event_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "kernel_ms": {k: round(v, 5) for k, v in kms.items()},
    }
    if name == "rgg22" and not args.no_learned:
        import learn

        cmp = learn.compare(csr, learn.load_params(args.params), seeds=seeds, graph=g)
        cmp["params"] = os.path.relpath(args.params, ROOT)
        line["learned_vs_plain"] = cmp
        log(f"learned vs plain: {json.dumps(cmp)}")
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        line["cpu_baseline"] = cpu_baseline(csr, seeds, r99)
        line["cpu_matchers"] = cpu_matchers(csr, seeds, r99)
        log(f"cpu matchers: {json.dumps(line['cpu_matchers'])}")
    g.close()
    return line


def run_batch(args) -> dict:
    t0 = time.perf_counter()
    specs = workloads.batch_sizes(args.batch_graphs, seed=args.seed)
    graphs = [workloads.batch_graph(*s) for s in specs]
    n_tot = sum(g.n for g in graphs)
    m_tot = sum(g.n_edges for g in graphs)
    log(f"generated {len(graphs)} graphs, {n_tot} vertices, {m_tot} edges in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    b = mmamd.Batch(graphs)
    log(f"batch upload {time.perf_counter() - t0:.1f} s")
    seeds = mmamd.reference_seeds(mmamd.MM_MAX_ROUNDS)
    res, per_graph, rounds = b.match(kind=mmamd.MM_WEIGHTED, maximal=True, seeds=seeds)
    log(f"batch maximal: rounds={res.rounds} (per graph {rounds.min()}..{rounds.max()}) match_ms={res.match_ms:.3f}")
    # every graph's result is the single-graph path's (a seeded sample of graphs, with the smallest and largest)
    rng = np.random.default_rng(args.seed)
    sizes = np.array([g.n for g in graphs])
    pick = sorted(set(rng.choice(len(graphs), size=min(args.batch_check, len(graphs)), replace=False).tolist()
                      + [int(np.argmin(sizes)), int(np.argmax(sizes))]))
    bad = 0
    for k in pick:
        single = mmamd.Graph(graphs[k]).match(kind=mmamd.MM_WEIGHTED, maximal=True, seeds=seeds)
        bad += int(not np.array_equal(single.match, per_graph[k]) or single.rounds != rounds[k])
    t0 = time.perf_counter()
    chk, gw, gs, gv = b.check()  # getWeight / testMatching of every graph, on the device
    check_ms = (time.perf_counter() - t0) * 1e3
    log(f"checked {len(pick)} graphs against the single-graph path: {bad} differ; device check {chk} "
        f"({int(gv.sum())} of {len(gv)} graphs valid) in {check_ms:.1f} ms")
    tune = b.autotune(kind=mmamd.MM_WEIGHTED, maximal=True, seeds=seeds) if args.tune else None
    tiles = tiles_of(tune) if tune else {"lanes": args.lanes, "block_size": args.block, "hub_degree": args.hub,
                                         "unroll": args.unroll}
    log(f"tiles {tiles} (autotune {tune})")

    def step():
        b.match(kind=mmamd.MM_WEIGHTED, maximal=True, seeds=seeds, fetch=False, **tiles)

    wall = timed(step, args.warmup, args.steps)
    ev = float(np.median([b.match(kind=mmamd.MM_WEIGHTED, maximal=True, seeds=seeds, fetch=False,
                                  **tiles)[0].match_ms for _ in range(5)]))
    tuned, per_tuned, _ = b.match(kind=mmamd.MM_WEIGHTED, maximal=True, seeds=seeds, **tiles)
    if not np.array_equal(tuned.match, res.match):
        raise RuntimeError("tuned tiles changed the batch result")
    return {
        "metric": "GTEPS, 4096-graph batch matched to convergence (weighted)",
        "value": round(m_tot / wall / 1e9, 3), "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall * 1e3, 4), "higher_is_better": True, "scaling": "strong",
        "vs_baseline": None, "dtype": "int32+f32cmp", "data": "synthetic (seeded, generated on the host)",
        "config": {"workload": f"batch{len(graphs)}: n log-uniform in [2^8, 2^16], even graphs 2-D grids "
                               f"(4-neighbour), odd graphs RGG r=0.55*sqrt(ln n/n), seed {args.seed}; type 9 "
                               f"rounds to convergence", "graphs": len(graphs), "n": n_tot, "m": m_tot,
                   "rounds": res.rounds, "rounds_per_graph": [int(rounds.min()), int(np.median(rounds)),
                                                              int(rounds.max())],
                   "tiles": tiles, "parallelism": "single GPU"},
        "device_event_ms": round(ev, 4), "gteps_device_events": round(m_tot / (ev * 1e-3) / 1e9, 3),
        "matched_vertices": chk["size"], "matching_weight": chk["weight"], "valid": bool(chk["valid"]),
        "graphs_valid": int(gv.sum()), "still_matchable": chk["matchable"], "device_check_ms": round(check_ms, 2),
        "graphs_checked_against_single_path": len(pick), "graphs_differing": bad, "autotune": tune,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=("rgg22", "rmat24", "batch4096"), default="rgg22")
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
    ap.add_argument("--no-cpu", action="store_true")
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

    t0 = time.perf_counter()
    if args.workload == "batch4096":
        line = run_batch(args)
    elif args.workload == "rmat24":
        csr = workloads.rmat(args.scale, seed=args.seed)
        log(f"generated R-MAT scale {args.scale}: n={csr.n} nnz={csr.nnz} in {time.perf_counter() - t0:.2f} s")
        line = run_single(args, csr, f"rmat{args.scale}", f"R-MAT scale {args.scale}, edge factor 16, "
                                                          f"(0.57, 0.19, 0.19, 0.05)")
    else:
        csr = workloads.rgg(args.log2n, seed=args.seed)
        log(f"generated RGG n={csr.n} nnz={csr.nnz} m={csr.n_edges} in {time.perf_counter() - t0:.2f} s")
        line = run_single(args, csr, f"rgg{args.log2n}", f"2^{args.log2n} RGG")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
