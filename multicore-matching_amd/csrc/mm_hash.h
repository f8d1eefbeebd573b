// The select hash of the matching rounds, shared by every kernel source.
#ifndef MM_HASH_H
#define MM_HASH_H

#include <hip/hip_runtime.h>

#include <cstdint>

__host__ __device__ inline uint32_t mm_rotl32(uint32_t x, uint32_t s) { return (x << s) | (x >> (32u - s)); }

// The select hash of src/matchgpu.cu:196-217: MD5-shaped, 16 steps, SHA-256
// constants K[0..15], MD5 round-1 rotations, per-step accumulation into h0..h3
// and g multiplied by the seed after every step.
__host__ __device__ inline uint32_t mm_select_hash(uint32_t i, uint32_t seed)
{
    uint32_t h0 = 0x67452301u, h1 = 0xefcdab89u, h2 = 0x98badcfeu, h3 = 0x10325476u;
    uint32_t a = h0, b = h1, c = h2, d = h3, g = i;
#define MM_STEP(K, R)                                         \
    {                                                         \
        const uint32_t f = (b & c) | (~b & d);                \
        const uint32_t t = d;                                 \
        d = c;                                                \
        c = b;                                                \
        b = b + mm_rotl32(a + f + (K) + g, (R));              \
        a = t;                                                \
        h0 += a; h1 += b; h2 += c; h3 += d;                   \
        g *= seed;                                            \
    }
    MM_STEP(0x428a2f98u, 7) MM_STEP(0x71374491u, 12) MM_STEP(0xb5c0fbcfu, 17) MM_STEP(0xe9b5dba5u, 22)
    MM_STEP(0x3956c25bu, 7) MM_STEP(0x59f111f1u, 12) MM_STEP(0x923f82a4u, 17) MM_STEP(0xab1c5ed5u, 22)
    MM_STEP(0xd807aa98u, 7) MM_STEP(0x12835b01u, 12) MM_STEP(0x243185beu, 17) MM_STEP(0x550c7dc3u, 22)
    MM_STEP(0x72be5d74u, 7) MM_STEP(0x80deb1feu, 12) MM_STEP(0x9bdc06a7u, 17) MM_STEP(0xc19bf174u, 22)
#undef MM_STEP
    return h0 + h1 + h2 + h3;
}

#endif
