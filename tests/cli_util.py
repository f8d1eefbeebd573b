"""Helpers of the match command-line driver tests: METIS files and output parsing."""
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MATCH = os.path.join(ROOT, "multicore-matching_amd", "match")


def distinct_weight_graph(n, avg_deg, seed):
    """Random simple graph with distinct float32 edge weights (no ties, so
    every matcher's result is independent of sort stability)."""
    import mmamd

    rng = np.random.default_rng(seed)
    m = n * avg_deg // 2
    a = rng.integers(0, n, 3 * m)
    b = rng.integers(0, n, 3 * m)
    keep = a != b
    e = np.unique(np.sort(np.stack([a[keep], b[keep]], 1), axis=1), axis=0)[:m]
    w = (rng.permutation(len(e)) + 1).astype(np.float32) / np.float32(1 << 20)
    return mmamd.CSR.from_edges(n, e, w)


def write_metis(path, csr):
    n = csr.n
    r = csr.ranges
    with open(path, "w") as f:
        f.write(f"% test graph\n{n} {csr.nnz // 2} 1\n")
        for i in range(n):
            nb = csr.neighbours[r[i, 0]:r[i, 1]]
            ww = csr.weights[r[i, 0]:r[i, 1]]
            f.write(" ".join(f"{int(u) + 1} {float(x):.9g}" for u, x in zip(nb, ww)) + "\n")


def run(args, timeout=300):
    p = subprocess.run([MATCH] + list(args), capture_output=True, text=True, timeout=timeout)
    return p.returncode, p.stdout, p.stderr


def parse(stdout):
    """{(type, threads): (size, weight)} and the header fields."""
    lines = [ln.split() for ln in stdout.strip().splitlines()]
    head = lines[0]
    rows = {}
    for ln in lines[1:]:
        # testMatching's report lines (-t) sit between the result lines
        if len(ln) != 10 or not ln[0].isdigit():
            continue
        rows[(int(ln[0]), int(ln[1]))] = (float(ln[2]), float(ln[4]), float(ln[6]), float(ln[8]))
    return head, rows
