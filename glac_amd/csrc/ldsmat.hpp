// ldsmat.hpp -- 3x3 complex matrices shared through LDS by the wavefronts of
// a block, one matrix per lane and slot.
//
// Lane-major: lane l's nine entries of slot s are consecutive complex pairs,
// so a wavefront writes or reads a matrix with nine 8-byte (FP32) or 16-byte
// (FP64) LDS accesses.  With an 18-word lane stride the 8-byte accesses of 16
// consecutive lanes fall on distinct bank pairs (18 l mod 32 runs over the 16
// even banks).
#pragma once
#include "lattice.hpp"

namespace glacx {

template <class T, int N> struct LdsMats {
    typename vec2<T>::type v[N][64][9];
};

template <class T, int N>
__device__ __forceinline__ void lds_put(LdsMats<T, N> &s, int slot, int lane, const M3<T> &m)
{
    using V2 = typename vec2<T>::type;
#pragma unroll
    for (int e = 0; e < 9; e++) {
        V2 z;
        z.x = m.e[e / 3][e % 3].r;
        z.y = m.e[e / 3][e % 3].i;
        s.v[slot][lane][e] = z;
    }
}

template <class T, int N> __device__ __forceinline__ M3<T> lds_get(const LdsMats<T, N> &s, int slot, int lane)
{
    M3<T> m;
#pragma unroll
    for (int e = 0; e < 9; e++) {
        const auto z = s.v[slot][lane][e];
        m.e[e / 3][e % 3] = {z.x, z.y};
    }
    return m;
}

} // namespace glacx
