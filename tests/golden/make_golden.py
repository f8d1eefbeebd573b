"""Writes the golden fixtures under tests/golden/ from the literal statement in
literal.py (pure Python, independent of the C oracle and of the HIP kernels).

  hash_vectors.json  (i, seed) -> h0+h1+h2+h3 of the select hash
  tiny_cases.json    tiny graphs -> final match array and rounds, for the four
                     reference matcher types and edge cases (unsorted rows,
                     duplicate entries, self-loops, isolated vertices, weights
                     <= -1, NaN and signed zeros, all-equal weights, extreme
                     barriers)

Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes
import json
import os
import random
import struct

import numpy as np

import literal

HERE = os.path.dirname(os.path.abspath(__file__))


def glibc_seeds(count, srand_seed=12345):
    libc = ctypes.CDLL("libc.so.6")
    libc.rand.restype = ctypes.c_int
    libc.srand(ctypes.c_uint(srand_seed))
    return [int(libc.rand()) for _ in range(count)]


def f32_bits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


def bits_f32(b):
    return struct.unpack("<f", struct.pack("<I", b))[0]


def random_graph(rng, n, p, *, unsorted=False, dups=False, loops=False, weight_mode="uniform"):
    rows = [[] for _ in range(n)]
    wts = [[] for _ in range(n)]
    for u in range(n):
        for v in range(u + 1, n):
            if rng.random() < p:
                if weight_mode == "uniform":
                    w = (rng.randrange(1 << 8) + 1) / 256.0  # coarse: ties happen
                elif weight_mode == "unit":
                    w = 1.0
                else:  # "nasty": signed zeros, <= -1, NaN, negatives
                    w = rng.choice([0.0, -0.0, -1.0, -2.5, -0.25, 0.5, 0.5, 1.0, float("nan"), 3.0])
                k = 2 if dups and rng.random() < 0.2 else 1
                for _ in range(k):
                    rows[u].append(v)
                    wts[u].append(w)
                    rows[v].append(u)
                    wts[v].append(w)
        if loops and rng.random() < 0.3:
            rows[u].append(u)
            wts[u].append(0.75)
    for u in range(n):
        order = list(range(len(rows[u])))
        if unsorted:
            rng.shuffle(order)
        else:
            order.sort(key=lambda j: rows[u][j])
        rows[u] = [rows[u][j] for j in order]
        wts[u] = [wts[u][j] for j in order]
    return rows, wts


def main():
    rng = random.Random(20111)
    seeds = glibc_seeds(255)

    vec = []
    hrng = random.Random(7)
    specials = [(0, 0), (0, 1), (0, seeds[0]), (1, 0), (1, 1), (12345, seeds[0]), (0xFFFFFFFF, 0xFFFFFFFF)]
    for i, s in specials + [(hrng.randrange(1 << 32), hrng.randrange(1 << 32)) for _ in range(400)]:
        vec.append([i, s, literal.select_hash(i, s)])
    with open(os.path.join(HERE, "hash_vectors.json"), "w") as f:
        json.dump({"source": "tests/golden/literal.py select_hash", "vectors": vec}, f)

    cases = []
    specs = []
    for n in (0, 1, 2, 3, 5, 8, 12, 16):
        for p in (0.3, 0.6):
            specs.append(dict(n=n, p=p))
    specs += [dict(n=10, p=0.5, unsorted=True), dict(n=12, p=0.4, unsorted=True, dups=True),
              dict(n=9, p=0.5, loops=True), dict(n=11, p=0.5, weight_mode="nasty"),
              dict(n=14, p=0.35, weight_mode="unit"), dict(n=16, p=0.25, unsorted=True, weight_mode="nasty"),
              dict(n=16, p=0.1), dict(n=6, p=1.0), dict(n=16, p=0.9, dups=True, loops=True)]
    barriers = [0x88B81733, 0, 0xFFFFFFFF, 0x40000000]
    for si, spec in enumerate(specs):
        rows, wts = random_graph(rng, spec["n"], spec["p"], unsorted=spec.get("unsorted", False),
                                 dups=spec.get("dups", False), loops=spec.get("loops", False),
                                 weight_mode=spec.get("weight_mode", "uniform"))
        n = spec["n"]
        offsets = [0]
        for r in rows:
            offsets.append(offsets[-1] + len(r))
        nbrs = [v for r in rows for v in r]
        wbits = [f32_bits(w) for r in wts for w in r]
        wf = [[bits_f32(f32_bits(w)) for w in r] for r in wts]
        for weighted in (False, True):
            for maximal in (False, True):
                barrier = barriers[(si + 2 * weighted + maximal) % len(barriers)] if si % 3 == 0 else 0x88B81733
                max_rounds = 255 if maximal else 20
                if si % 5 == 4:
                    max_rounds = 3 if not maximal else 4  # truncated runs keep colours in the output
                m, R = literal.run(n, rows, wf, weighted, maximal, seeds, barrier=barrier, max_rounds=max_rounds)
                cases.append(dict(spec=spec, n=n, offsets=offsets, neighbours=nbrs, weight_bits=wbits,
                                  weighted=weighted, maximal=maximal, max_rounds=max_rounds, barrier=barrier,
                                  match=m, rounds=R))
    with open(os.path.join(HERE, "tiny_cases.json"), "w") as f:
        json.dump({"source": "tests/golden/literal.py run", "srand": 12345, "seeds": seeds, "cases": cases}, f)
    print(f"wrote {len(vec)} hash vectors and {len(cases)} tiny cases")


if __name__ == "__main__":
    main()
