"""Device memory of graph and batch handles is returned when they are
destroyed: repeated construct/destroy cycles leave the device's free memory
where it was (every array a handle allocates is released)."""
import ctypes

import numpy as np
import pytest

import mmamd

pytestmark = pytest.mark.gpu


def free_bytes() -> int:
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return int(free.value)


def test_handles_release_their_device_memory():
    rgg = mmamd.gen_rgg(1 << 20, seed=51, weighted=True)       # 32-bit keys, packed rows
    rmat = mmamd.gen_rmat(16, 8, seed=52, weighted=True)      # vertices without neighbours settled by the select
    batch = [mmamd.gen_rgg(n, seed=n, weighted=True) for n in (300, 5000, 20000)]
    seeds = mmamd.reference_seeds(mmamd.MM_MAX_ROUNDS)

    def cycle():
        for csr in (rgg, rmat):
            g = mmamd.Graph(csr)
            g.match(kind=mmamd.MM_WEIGHTED, maximal=True, seeds=seeds)
            g.check()
            g.match_learned(None, kind=mmamd.MM_WEIGHTED, maximal=True, seeds=seeds)
            g.close()
        b = mmamd.Batch(batch)
        b.match(kind=mmamd.MM_WEIGHTED, maximal=True, seeds=seeds)
        b.close()

    cycle()  # (first use: the runtime's own allocations)
    before = free_bytes()
    for _ in range(3):
        cycle()
    after = free_bytes()
    # one RGG handle holds about 600 MB; a leaked array of it would show up three times
    assert before - after < 32 << 20, (before, after)
