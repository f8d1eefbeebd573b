"""Training and evaluation of the learned selection (include/mm_learn.h).

No counterpart in the reference (BasFaggingerAuer/Multicore-Matching): its
select step colours every active vertex blue with one fixed probability
(gSelect, src/matchgpu.cu:184-218).  Here the blue probability comes from an
8 -> 32 -> 32 -> 1 network, trained on round traces of the plain rounds:

  samples  one per vertex matched in a plain round: its features at that
           round's select, label 1 if it was blue, weighted by the weight of
           the edge it was matched along (only matched vertices teach);
  target   the network learns P(blue | matched, features): it moves each
           vertex's colour towards the one that, in the plain rounds, ended
           in heavy matches;
  training minibatch Adam on the weighted logistic loss, by the device
           forward/backward kernels (mm_mlp_train).

Trace collection, training and the matchings all run in the HIP kernels of
libmmamd.so; this module only drives them.

Command line:
  python learn.py train --out mlp_params.json [--graphs 200] [--min-log2n 16] [--max-log2n 20]
  python learn.py eval --params mlp_params.json [--log2n 22]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import mmamd  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_PARAMS = os.path.join(HERE, "mlp_params.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def training_graphs(count: int = 200, min_log2n: int = 16, max_log2n: int = 20, seed: int = 2024):
    """Seeded training set: graph k is an RGG (even k) or an R-MAT graph (odd
    k, edge factor 16, a, b, c = 0.57, 0.19, 0.19), n = 2^e with e uniform in
    [min_log2n, max_log2n] (log-uniform sizes); float32 weights on every edge."""
    rng = np.random.default_rng(seed)
    for k in range(count):
        e = float(rng.uniform(min_log2n, max_log2n))
        gseed = int(rng.integers(1, 2**62))
        if k % 2 == 0:
            yield k, mmamd.gen_rgg(int(round(2.0 ** e)), seed=gseed, weighted=True)
        else:
            yield k, mmamd.gen_rmat(int(round(e)), seed=gseed, weighted=True)


def collect_traces(graphs, samples_per_graph: int = 16384, seeds=None, device: int = 0, progress: bool = False):
    """Plain-round traces on the device: (x[N, 8], y[N], w[N]) over the graphs."""
    if seeds is None:
        seeds = mmamd.reference_seeds(mmamd.MM_MAX_ROUNDS)
    xs, ys, ws = [], [], []
    t0 = time.perf_counter()
    for k, csr in graphs:
        if progress and k % 10 == 0:
            log(f"trace graph {k}: n={csr.n} nnz={csr.nnz} ({time.perf_counter() - t0:.1f} s)")
        g = mmamd.Graph(csr, device=device)
        keep = int(min(65536, math.ceil(65536.0 * samples_per_graph / max(1.0, 0.9 * csr.n))))
        _, s = g.match_learned(None, kind=mmamd.MM_WEIGHTED, maximal=True, seeds=seeds, trace=True,
                               trace_keep=keep, trace_cap=2 * samples_per_graph + 1024)
        g.close()
        xs.append(s["x"])
        ys.append(s["y"])
        ws.append(s["w"])
    return np.concatenate(xs), np.concatenate(ys), np.concatenate(ws)


def train(x, y, w, init_seed: int = 1, epochs: int = 4, batch: int = 4096, lr: float = 3e-3, seed: int = 7,
          device: int = 0):
    """Device training from a He-uniform start: (params, per-epoch loss)."""
    p0 = mmamd.mlp_init(init_seed)
    return mmamd.mlp_train(x, y, w, p0, epochs=epochs, batch=batch, lr=lr, seed=seed, device=device)


def save_params(path: str, params: np.ndarray, meta: dict):
    rec = dict(meta)
    rec["layout"] = "W1[32][8], b1[32], W2[32][32], b2[32], W3[32], b3 (include/mm_learn.h)"
    rec["params_f32_bits"] = [int(b) for b in np.asarray(params, np.float32).view(np.uint32)]
    with open(path, "w") as f:
        json.dump(rec, f, indent=1)


def load_params(path: str = DEFAULT_PARAMS) -> np.ndarray:
    with open(path) as f:
        rec = json.load(f)
    p = np.array(rec["params_f32_bits"], dtype=np.uint32).view(np.float32)
    if p.shape != (mmamd.MM_MLP_PARAMS,):
        raise ValueError(f"{path}: expected {mmamd.MM_MLP_PARAMS} parameters")
    return p


def compare(csr, params, seeds=None, graph=None, max_rounds: int = 0, reps: int = 5) -> dict:
    """Plain and learned maximal matchings of one graph on the device:
    rounds, matched vertices, weight, device time (result left in HBM), and the first round at
    which the learned run's weight reaches the plain run's final weight."""
    if seeds is None:
        seeds = mmamd.reference_seeds(mmamd.MM_MAX_ROUNDS)
    g = graph if graph is not None else mmamd.Graph(csr)
    # untimed first runs: buffers and first kernel launches
    g.match(kind=mmamd.MM_WEIGHTED, maximal=True, seeds=seeds, max_rounds=max_rounds, fetch=False)
    g.match_learned(params, kind=mmamd.MM_WEIGHTED, maximal=True, seeds=seeds, max_rounds=max_rounds, fetch=False)
    # timed runs leave the result on the device (match_ms is the device time of
    # the rounds and the unpermute); the arrays are copied out afterwards.  A
    # maximal run's time includes its host counter read-backs, so each time is
    # the median of `reps` runs.
    def median_run(fn):
        runs = [fn() for _ in range(reps)]
        runs.sort(key=lambda x: x.match_ms)
        return runs[len(runs) // 2]
    plain = median_run(lambda: g.match(kind=mmamd.MM_WEIGHTED, maximal=True, seeds=seeds, max_rounds=max_rounds,
                                       fetch=False))
    plain_match = g.copy_match()
    learned = median_run(lambda: g.match_learned(params, kind=mmamd.MM_WEIGHTED, maximal=True, seeds=seeds,
                                                 max_rounds=max_rounds, fetch=False))
    learned_match = g.copy_match()
    wp, sp = mmamd.get_weight(csr, plain_match)
    wl, sl = mmamd.get_weight(csr, learned_match)
    reach, reach_ms = None, None
    for r in range(1, learned.rounds + 1):
        g.match_learned(params, kind=mmamd.MM_WEIGHTED, maximal=False, seeds=seeds, max_rounds=r, fetch=False)
        if mmamd.get_weight(csr, g.copy_match())[0] >= wp:
            m = median_run(lambda: g.match_learned(params, kind=mmamd.MM_WEIGHTED, maximal=False, seeds=seeds,
                                                   max_rounds=r, fetch=False))
            reach, reach_ms = r, m.match_ms
            break
    # the same kernels with the plain colouring: the cost of the learned rounds without the network
    through = median_run(lambda: g.match_learned(None, kind=mmamd.MM_WEIGHTED, maximal=True, seeds=seeds,
                                                 max_rounds=max_rounds, fetch=False))
    out = {
        "plain": {"rounds": plain.rounds, "matched": sp, "weight": wp, "ms": plain.match_ms,
                  "valid": mmamd.test_matching(csr, plain_match)[0]},
        "learned": {"rounds": learned.rounds, "matched": sl, "weight": wl, "ms": learned.match_ms,
                    "valid": mmamd.test_matching(csr, learned_match)[0],
                    "rounds_to_plain_weight": reach, "ms_to_plain_weight": reach_ms},
        "plain_colouring_in_learned_kernels_ms": through.match_ms,
        "weight_ratio": wl / wp if wp else None,
    }
    if graph is None:
        g.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--out", default=DEFAULT_PARAMS)
    t.add_argument("--graphs", type=int, default=200)
    t.add_argument("--min-log2n", type=int, default=16)
    t.add_argument("--max-log2n", type=int, default=20)
    t.add_argument("--samples-per-graph", type=int, default=16384)
    t.add_argument("--epochs", type=int, default=4)
    t.add_argument("--seed", type=int, default=2024)
    e = sub.add_parser("eval")
    e.add_argument("--params", default=DEFAULT_PARAMS)
    e.add_argument("--log2n", type=int, default=22)
    args = ap.parse_args()
    if args.cmd == "train":
        t0 = time.perf_counter()
        x, y, w = collect_traces(training_graphs(args.graphs, args.min_log2n, args.max_log2n, args.seed),
                                 args.samples_per_graph, progress=True)
        t1 = time.perf_counter()
        log(f"traces: {len(x)} samples from {args.graphs} graphs in {t1 - t0:.1f} s (blue fraction {y.mean():.3f})")
        p, losses = train(x, y, w, epochs=args.epochs)
        t2 = time.perf_counter()
        log(f"training: {args.epochs} epochs in {t2 - t1:.2f} s, loss per epoch {np.round(losses, 5).tolist()}")
        save_params(args.out, p, {"graphs": args.graphs, "log2n": [args.min_log2n, args.max_log2n],
                                  "seed": args.seed, "samples": int(len(x)), "epochs": args.epochs,
                                  "epoch_loss": [float(v) for v in losses]})
        print(json.dumps({"samples": int(len(x)), "epoch_loss": [float(v) for v in losses],
                          "trace_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2), "out": args.out}))
    else:
        p = load_params(args.params)
        csr = mmamd.gen_rgg(1 << args.log2n, seed=12345, weighted=True)
        print(json.dumps(compare(csr, p)))


if __name__ == "__main__":
    main()
