// rng.hpp -- counter-based per-link random streams.
//
// One Philox4x32-10 block (Salmon, Moraes, Dror, Shaw, SC'11), keyed by the
// 64-bit run seed with counter (lex site, mu | tag << 8, lo32(sweep),
// hi32(sweep)), seeds a xoshiro128** state (Blackman & Vigna) that the link
// then draws from.  Every link of every sweep owns an independent stream, so
// the result does not depend on the launch shape or on the order in which
// threads run.  GLAC's mt19937_64 streams (GLAC/system.cpp:40-41,
// GLAC/math/matrices/su3matrixgenerator.cpp:22-24) are inherently sequential
// and are replaced by this definition.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace glacx {

enum StreamTag : uint32_t { STREAM_METROPOLIS = 1, STREAM_HOT_START = 2, STREAM_LEARNED = 3, STREAM_TRAIN = 4 };

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        const uint32_t n3 = (uint32_t)p0;
        c[0] = n0;
        c[1] = n1;
        c[2] = n2;
        c[3] = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

struct Stream {
    uint32_t s[4];

    __device__ __forceinline__ Stream(uint64_t seed, uint32_t site, uint32_t mu, uint32_t tag, uint64_t sweep)
    {
        s[0] = site;
        s[1] = mu | (tag << 8);
        s[2] = (uint32_t)sweep;
        s[3] = (uint32_t)(sweep >> 32);
        philox4x32_10(s, (uint32_t)seed, (uint32_t)(seed >> 32));
        if ((s[0] | s[1] | s[2] | s[3]) == 0u) s[0] = 1u;
    }

    __device__ __forceinline__ uint32_t next()
    {
        const uint32_t x = s[1] * 5u;
        const uint32_t result = ((x << 7) | (x >> 25)) * 9u;
        const uint32_t t = s[1] << 9;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = (s[3] << 11) | (s[3] >> 21);
        return result;
    }

    // two uniforms on the 2^16 midpoints of [-1/2, 1/2) from one 32-bit
    // output, lo from its low and hi from its high 16 bits:
    // (k + 1/2) 2^-16 - 1/2 = (2k - 2^16 + 1) 2^-17, exact in float and double
    // and never zero
    template <class T> __device__ __forceinline__ void centered_pair(T &lo, T &hi)
    {
        const uint32_t b = next();
        lo = T(b & 0xFFFFu) * T(0x1p-16) - T(0.5 - 0x1p-17);
        hi = T(b >> 16) * T(0x1p-16) - T(0.5 - 0x1p-17);
    }

    // uniform in (-1, 1) with 32 bits, in double: (k + 1/2) 2^-31 - 1
    __device__ __forceinline__ double symmetric() { return ((double)next() + 0.5) * 0x1p-31 - 1.0; }

    // uniform in (0, 1) with 32 bits, in double
    __device__ __forceinline__ double open() { return ((double)next() + 0.5) * 0x1p-32; }
};

} // namespace glacx
