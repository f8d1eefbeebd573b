"""Literal pure-Python statement of the reference's matching rounds, for tiny inputs.

Written independently of oracle/mm_oracle.c, phase by phase from the text of
src/matchgpu.cu (BasFaggingerAuer/Multicore-Matching), so that the two can pin
each other.  Its hash constants are not typed in: they are re-derived from
their mathematical definitions —
  * K[j]: the first 32 bits of the fractional part of the cube root of the
    j-th prime (the SHA-256 round constants; src/matchgpu.cu:167-169 uses
    K[0..15]);
  * the MD5 initial words A..D from the byte strings of RFC 1321 §3.3
    (01 23 45 67, 89 ab cd ef, fe dc ba 98, 76 54 32 10, little-endian);
  * the MD5 round-1 rotations 7, 12, 17, 22 (RFC 1321 §3.4, src/matchgpu.cu:177).
"""
from __future__ import annotations

MASK = 0xFFFFFFFF


def _primes(count):
    out, k = [], 2
    while len(out) < count:
        if all(k % p for p in out if p * p <= k):
            out.append(k)
        k += 1
    return out


def _icbrt(x: int) -> int:
    """floor(cube root of x) for a non-negative integer."""
    lo, hi = 0, 1 << ((x.bit_length() + 2) // 3 + 1)
    while lo < hi:
        mid = (lo + hi + 1) // 2
        if mid ** 3 <= x:
            lo = mid
        else:
            hi = mid - 1
    return lo


def sha256_k(count=16):
    # floor(cbrt(p) * 2^32) mod 2^32 == floor(cbrt(p * 2^96)) mod 2^32
    return [_icbrt(p << 96) & MASK for p in _primes(count)]


def md5_iv():
    words = [bytes([0x01, 0x23, 0x45, 0x67]), bytes([0x89, 0xAB, 0xCD, 0xEF]),
             bytes([0xFE, 0xDC, 0xBA, 0x98]), bytes([0x76, 0x54, 0x32, 0x10])]
    return [int.from_bytes(w, "little") for w in words]


K = sha256_k(16)
IV = md5_iv()
ROT = [7, 12, 17, 22]


def rotl(x, s):
    return ((x << s) | (x >> (32 - s))) & MASK


def select_hash(i: int, seed: int) -> int:
    """h0+h1+h2+h3 of src/matchgpu.cu:196-217."""
    h = list(IV)
    a, b, c, d = h
    g = i & MASK
    for j in range(16):
        f = (b & c) | ((~b & MASK) & d)
        e = d
        d = c
        c = b
        b = (b + rotl((a + f + K[j] + g) & MASK, ROT[j % 4])) & MASK
        a = e
        h[0] = (h[0] + a) & MASK
        h[1] = (h[1] + b) & MASK
        h[2] = (h[2] + c) & MASK
        h[3] = (h[3] + d) & MASK
        g = (g * seed) & MASK
    return sum(h) & MASK


def run(n, rows, weights, weighted, maximal, seeds, barrier=0x88B81733, max_rounds=None):
    """rows[i] = list of neighbour ids in CSR order, weights[i] = matching list
    of float32-representable weights.  Returns (match, rounds)."""
    import numpy as np

    f32 = np.float32
    if max_rounds is None:
        max_rounds = 255 if maximal else 20
    match = [0] * n
    requests = [0] * n
    count = 0
    keep = True
    while (keep if maximal else True) and count < max_rounds:
        seed = int(seeds[count])
        count += 1
        keep = False
        # gSelect / gaSelect (src/matchgpu.cu:184-258)
        for i in range(n):
            if match[i] >= 2:
                continue
            keep = True
            match[i] = 0 if select_hash(i, seed) < barrier else 1
        # grRequest / gwRequest (src/matchgpu.cu:286-327, 360-409)
        for i in range(n):
            if match[i] == 0:
                dead = 1
                candidate = n
                max_w = f32(-1.0)
                done = False
                for j, ni in enumerate(rows[i]):
                    nm = match[ni]
                    if nm < 4:
                        if nm == 1:
                            if not weighted:
                                requests[i] = ni
                                done = True
                                break
                            nw = f32(weights[i][j])
                            if nw > max_w:
                                max_w = nw
                                candidate = ni
                        dead = 0
                if not done:
                    requests[i] = candidate + dead if weighted else n + dead
            else:
                requests[i] = n
        # grRespond / gwRespond (src/matchgpu.cu:329-357, 411-451)
        snapshot = list(requests)
        for i in range(n):
            if match[i] != 1:
                continue
            candidate = n
            max_w = f32(-1.0)
            for j, ni in enumerate(rows[i]):
                if match[ni] == 0 and snapshot[ni] == i:
                    if not weighted:
                        candidate = ni
                        break
                    nw = f32(weights[i][j])
                    if nw > max_w:
                        max_w = nw
                        candidate = ni
            if candidate < n:
                requests[i] = candidate
        # gMatch (src/matchgpu.cu:260-283)
        final = list(requests)
        for i in range(n):
            r = final[i]
            if r == n + 1:
                match[i] = 2
            elif r < n and final[r] == i:
                match[i] = 4 + min(i, r)
    return match, count
