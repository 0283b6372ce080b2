// lattice.hpp -- device layout of the gauge field and periodic neighbour
// arithmetic.
//
// Sites are stored parity-split: idx = p * Vh + (lex >> 1), with
// lex = x + Nx (y + Ny (z + Nz t)) GLAC's site index
// (GLAC/parallelization/index.h:32-35), p = (x + y + z + t) & 1 and Vh = V / 2.
// Every extent is even, so within an x-row the sites of one parity are
// x0, x0 + 2, ... and lex >> 1 is unique per parity.  A checkerboard
// sub-sweep therefore touches one contiguous half of every plane.
//
// A field is 36 complex planes (mu-major, then matrix entry row-major), each
// V entries long: entry (i, j) of U_mu at idx lives at
// plane[(mu * 9 + 3 i + j) * V + idx].  One wavefront reading one plane for 64
// consecutive idx issues one contiguous 512 B (FP32) / 1 KiB (FP64) access.
//
// GLAC's shift() and Communicator link getters (GLAC/math/lattice.h:824-1144,
// GLAC/parallelization/communicator.cpp:83-258) become the index arithmetic
// below: the whole lattice is resident on one GPU.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "su3.hpp"

namespace glacx {

struct Geom {
    int n[4];      // Nx, Ny, Nz, Nt
    int nxh;       // Nx / 2
    int vs_h;      // Nx Ny Nz / 2: sites of one parity in one time slice
    int vh;        // V / 2
    int v;         // V
};

struct Coord {
    int c[4];
};

// (parity-split idx) -> coordinates
__device__ __forceinline__ Coord decode(const Geom &g, int idx)
{
    const int p = idx >= g.vh;
    int h = idx - p * g.vh;
    Coord r;
    const int hx = h % g.nxh;
    h /= g.nxh;
    r.c[1] = h % g.n[1];
    h /= g.n[1];
    r.c[2] = h % g.n[2];
    r.c[3] = h / g.n[2];
    r.c[0] = 2 * hx + ((r.c[1] + r.c[2] + r.c[3] + p) & 1);
    return r;
}

__device__ __forceinline__ int lex_of(const Geom &g, const Coord &r)
{
    return r.c[0] + g.n[0] * (r.c[1] + g.n[1] * (r.c[2] + g.n[2] * r.c[3]));
}

__device__ __forceinline__ int index_of(const Geom &g, const Coord &r)
{
    const int p = (r.c[0] + r.c[1] + r.c[2] + r.c[3]) & 1;
    return p * g.vh + (lex_of(g, r) >> 1);
}

// coordinates of r + s * e_mu, periodic
__device__ __forceinline__ Coord step(const Geom &g, Coord r, int mu, int s)
{
    int v = r.c[mu] + s;
    const int N = g.n[mu];
    v = v < 0 ? v + N : (v >= N ? v - N : v);
    r.c[mu] = v;
    return r;
}

// ---- field planes --------------------------------------------------------
template <class T> struct vec2;
template <> struct vec2<float> {
    using type = float2;
};
template <> struct vec2<double> {
    using type = double2;
};

template <class T>
__device__ __forceinline__ M3<T> load_link(const typename vec2<T>::type *__restrict__ f, int v, int mu, int idx)
{
    M3<T> m;
    const typename vec2<T>::type *p = f + (size_t)mu * 9 * v + idx;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const typename vec2<T>::type z = p[(size_t)k * v];
        m.e[k / 3][k % 3] = {z.x, z.y};
    }
    return m;
}

template <class T>
__device__ __forceinline__ void store_link(typename vec2<T>::type *__restrict__ f, int v, int mu, int idx, const M3<T> &m)
{
    typename vec2<T>::type *p = f + (size_t)mu * 9 * v + idx;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        typename vec2<T>::type z;
        z.x = m.e[k / 3][k % 3].r;
        z.y = m.e[k / 3][k % 3].i;
        p[(size_t)k * v] = z;
    }
}

// Wilson staple A_mu(x) = sum_{nu != mu} [ U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
//                                        + U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu) ]
// (GLAC/actions/wilsongaugeaction.cpp:29-53).
template <class T>
__device__ __forceinline__ M3<T> staple(const typename vec2<T>::type *__restrict__ f, const Geom &g, const Coord &x,
                                        int mu)
{
    M3<T> a = zero3<T>();
    const Coord xpm = step(g, x, mu, +1);
#pragma unroll 1
    for (int d = 1; d < 4; d++) {
        const int nu = (mu + d) & 3;
        const Coord xpn = step(g, x, nu, +1);
        const Coord xmn = step(g, x, nu, -1);
        const Coord xpmmn = step(g, xpm, nu, -1);
        {
            const M3<T> u1 = load_link<T>(f, g.v, nu, index_of(g, xpm));
            const M3<T> u2 = load_link<T>(f, g.v, mu, index_of(g, xpn));
            const M3<T> u3 = load_link<T>(f, g.v, nu, index_of(g, x));
            add_to(a, mul_ad(mul_ad(u1, u2), u3));
        }
        {
            const M3<T> u1 = load_link<T>(f, g.v, nu, index_of(g, xpmmn));
            const M3<T> u2 = load_link<T>(f, g.v, mu, index_of(g, xmn));
            const M3<T> u3 = load_link<T>(f, g.v, nu, index_of(g, xmn));
            add_to(a, mul(mul_dd(u1, u2), u3));
        }
    }
    return a;
}

} // namespace glacx
