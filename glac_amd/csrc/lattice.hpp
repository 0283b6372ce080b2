// lattice.hpp -- device layout of the gauge field and periodic neighbour
// arithmetic.
//
// Sites are stored in 4D tiles.  Each extent n_k is split into tiles of
// b_k = 4 sites (2 when n_k is not a multiple of 4; every extent is even), so
// a tile holds b0 b1 b2 b3 sites (a power of two).  Tiles are ordered
// lexicographically (x fastest) and each tile stores its even sites, then its
// odd sites (parity p = (x + y + z + t) & 1), each half in intra-tile
// lexicographic order halved:
//     idx = tile * TV + p * TV / 2 + (intra_lex >> 1),
//     tile = tx + T0 (ty + T1 (tz + T2 tt)),  intra_lex = lx + b0 (ly + b1 (lz + b2 lt)).
// Since b0 is even, intra_lex >> 1 is unique per parity.
//
// Why tiles: a 4D stencil touches the 8 nearest neighbours of every site.
// With rows of full x-extent the t-neighbour of a site lies a whole 3D slice
// away (9.4 MB of FP32 links at 32^4, more than an XCD's 4 MB L2), so every
// link is fetched from beyond L2 several times per sweep.  In a 4^4 tile most
// neighbours are in the same tile; only the tile faces (a quarter of the
// sites per direction) reach into the next tile.  Sites of one parity within
// a tile are contiguous, so a checkerboard sub-sweep and a 64-lane wavefront
// still read contiguous runs.
//
// A field is 36 complex planes (mu-major, then matrix entry row-major), each
// V entries long: entry (i, j) of U_mu at idx lives at
// plane[(mu * 9 + 3 i + j) * V + idx].
//
// GLAC's shift() and Communicator link getters (GLAC/math/lattice.h:824-1144,
// GLAC/parallelization/communicator.cpp:83-258) become the index arithmetic
// below: the whole lattice is resident on one GPU.  GLAC's own site index
// lex = x + Nx (y + Ny (z + Nz t)) (GLAC/parallelization/index.h:32-35) is
// kept for host transfers and as the random-stream counter.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "su3.hpp"

namespace glacx {

// Division by a small runtime divisor d < 2^10 for dividends n < 2^30:
// q = (n * m) >> 40 with m = floor(2^40 / d) + 1 is exact because
// n * (m - 2^40 / d) / 2^40 < n / 2^40 < 1 / d.
struct FastDiv {
    uint64_t m;
    int d;
};

__host__ __device__ inline FastDiv make_fastdiv(int d)
{
    FastDiv f;
    f.d = d;
    f.m = ((uint64_t)1 << 40) / (uint64_t)d + 1;
    return f;
}

__host__ __device__ inline int fdiv(int n, const FastDiv &f) { return (int)(((uint64_t)n * f.m) >> 40); }

struct Geom {
    int n[4];       // Nx, Ny, Nz, Nt
    int lb[4];      // log2 of the tile extents b_k (1 or 2)
    int tn[4];      // tile grid extents n_k / b_k
    int ltv;        // log2 of the tile volume TV
    int v;          // V
    int vh;         // V / 2
    int vs;         // sites per time slice, Nx Ny Nz
    FastDiv d_t0, d_t1, d_t2;
};

__host__ inline Geom make_geom(const int dims[4])
{
    Geom g;
    g.ltv = 0;
    for (int k = 0; k < 4; k++) {
        g.n[k] = dims[k];
        g.lb[k] = (dims[k] % 4 == 0) ? 2 : 1;
        g.tn[k] = dims[k] >> g.lb[k];
        g.ltv += g.lb[k];
    }
    g.v = dims[0] * dims[1] * dims[2] * dims[3];
    g.vh = g.v / 2;
    g.vs = dims[0] * dims[1] * dims[2];
    g.d_t0 = make_fastdiv(g.tn[0]);
    g.d_t1 = make_fastdiv(g.tn[1]);
    g.d_t2 = make_fastdiv(g.tn[2]);
    return g;
}

// Blocks are handed to the 8 XCDs (each with its own L2) round-robin in
// launch order.  Remap so that XCD k processes the k-th contiguous eighth of
// the block range: neighbouring sites then share an L2.
__device__ __forceinline__ int xcd_block(int b, int nb)
{
#ifdef GLACX_NO_XCD_REMAP
    return b;
#endif
    const int xcd = b & 7, local = b >> 3;
    const int per = nb >> 3, rem = nb & 7;
    return xcd * per + (xcd < rem ? xcd : rem) + local;
}

struct Coord {
    int c[4];
};

// storage idx -> coordinates
__host__ __device__ inline Coord decode(const Geom &g, int idx)
{
    const int tile = idx >> g.ltv;
    const int within = idx & ((1 << g.ltv) - 1);
    const int p = within >> (g.ltv - 1);
    int j = within & ((1 << (g.ltv - 1)) - 1);
    // j = hx + (b0 / 2) (ly + b1 (lz + b2 lt))
    const int hx = j & ((1 << (g.lb[0] - 1)) - 1);
    j >>= g.lb[0] - 1;
    const int ly = j & ((1 << g.lb[1]) - 1);
    j >>= g.lb[1];
    const int lz = j & ((1 << g.lb[2]) - 1);
    const int lt = j >> g.lb[2];
    const int r0 = fdiv(tile, g.d_t0);
    const int tx = tile - r0 * g.tn[0];
    const int r1 = fdiv(r0, g.d_t1);
    const int ty = r0 - r1 * g.tn[1];
    const int tt = fdiv(r1, g.d_t2);
    const int tz = r1 - tt * g.tn[2];
    Coord r;
    r.c[1] = (ty << g.lb[1]) + ly;
    r.c[2] = (tz << g.lb[2]) + lz;
    r.c[3] = (tt << g.lb[3]) + lt;
    // tile origins are even, so the parity of the site is that of its offsets
    r.c[0] = (tx << g.lb[0]) + 2 * hx + ((ly + lz + lt + p) & 1);
    return r;
}

__host__ __device__ inline int lex_of(const Geom &g, const Coord &r)
{
    return r.c[0] + g.n[0] * (r.c[1] + g.n[1] * (r.c[2] + g.n[2] * r.c[3]));
}

__host__ __device__ inline int index_of(const Geom &g, const Coord &r)
{
    const int p = (r.c[0] + r.c[1] + r.c[2] + r.c[3]) & 1;
    const int tile = (r.c[0] >> g.lb[0]) +
                     g.tn[0] * ((r.c[1] >> g.lb[1]) + g.tn[1] * ((r.c[2] >> g.lb[2]) + g.tn[2] * (r.c[3] >> g.lb[3])));
    const int intra = (r.c[0] & ((1 << g.lb[0]) - 1)) +
                      ((r.c[1] & ((1 << g.lb[1]) - 1)) +
                       ((r.c[2] & ((1 << g.lb[2]) - 1)) + ((r.c[3] & ((1 << g.lb[3]) - 1)) << g.lb[2]) << g.lb[1])
                       << g.lb[0]);
    return (tile << g.ltv) + (p << (g.ltv - 1)) + (intra >> 1);
}

// storage idx of the h-th site of parity p (h < V / 2): tiles in order, each
// contributing its TV / 2 sites of that parity.
__host__ __device__ inline int parity_index(const Geom &g, int p, int h)
{
    const int tile = h >> (g.ltv - 1);
    return (tile << g.ltv) + (p << (g.ltv - 1)) + (h & ((1 << (g.ltv - 1)) - 1));
}

// storage idx of the r-th site of time slice t (r < Nx Ny Nz): the tiles of
// tile row tt = t / b3 in order, each contributing its b0 b1 b2 sites at
// local time lt = t % b3 (even then odd parity, each a contiguous run).
__host__ __device__ inline int slice_index(const Geom &g, int t, int r)
{
    const int l3 = g.lb[0] + g.lb[1] + g.lb[2];
    const int tile = (r >> l3) + (g.tn[0] * g.tn[1] * g.tn[2]) * (t >> g.lb[3]);
    const int rin = r & ((1 << l3) - 1);
    const int p = rin >> (l3 - 1);
    const int jj = rin & ((1 << (l3 - 1)) - 1);
    const int lt = t & ((1 << g.lb[3]) - 1);
    return (tile << g.ltv) + (p << (g.ltv - 1)) + (lt << (l3 - 1)) + jj;
}

// coordinates of r + s * e_mu (s = +-1), periodic.  Written with selects over
// the four components so that a runtime mu never indexes a private array.
__host__ __device__ inline Coord step(const Geom &g, Coord r, int mu, int s)
{
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int v = r.c[k] + s;
        v = v < 0 ? g.n[k] - 1 : (v == g.n[k] ? 0 : v);
        r.c[k] = k == mu ? v : r.c[k];
    }
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

// Field access.  The 9 complex planes of one direction mu are addressed
// through one buffer resource (base = plane (mu, 0), range 9 V entries, in
// scalar registers when mu is wave-uniform); a link load is then 9
// buffer_load_dwordx2/x4 with the per-lane byte offset idx * sizeof(V2) in a
// VGPR and the plane offset k V sizeof(V2) in an SGPR -- no per-load address
// arithmetic.  The context limits 9 V sizeof(V2) to below 2^32 bytes.
constexpr int kBufferFlags = 0x00020000;  // gfx9 raw buffer: 32-bit data format, no swizzle

template <class T> struct DirView {
    __amdgpu_buffer_rsrc_t r;
    int plane_bytes;
};

template <class T>
__device__ __forceinline__ DirView<T> dir_view(const typename vec2<T>::type *f, int v, int mu)
{
    using V2 = typename vec2<T>::type;
    DirView<T> d;
    const unsigned range = 9u * (unsigned)v * (unsigned)sizeof(V2);
    d.r = __builtin_amdgcn_make_buffer_rsrc((void *)(f + (size_t)mu * 9 * v), (short)0, (int)range, kBufferFlags);
    d.plane_bytes = v * (int)sizeof(V2);
    return d;
}

__device__ __forceinline__ float2 buf_load(const DirView<float> &d, unsigned off, int k)
{
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(d.r, off, k * d.plane_bytes, 0));
}
__device__ __forceinline__ double2 buf_load(const DirView<double> &d, unsigned off, int k)
{
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(d.r, off, k * d.plane_bytes, 0));
}
__device__ __forceinline__ void buf_store(const DirView<float> &d, unsigned off, int k, float2 z)
{
    using U2 = unsigned int __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, z), d.r, off, k * d.plane_bytes, 0);
}
__device__ __forceinline__ void buf_store(const DirView<double> &d, unsigned off, int k, double2 z)
{
    using U4 = unsigned int __attribute__((ext_vector_type(4)));
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(U4, z), d.r, off, k * d.plane_bytes, 0);
}

template <class T> __device__ __forceinline__ M3<T> load_link(const DirView<T> &d, int idx)
{
    using V2 = typename vec2<T>::type;
    const unsigned off = (unsigned)idx * (unsigned)sizeof(V2);
    M3<T> m;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const V2 z = buf_load(d, off, k);
        m.e[k / 3][k % 3] = {z.x, z.y};
    }
    return m;
}

template <class T> __device__ __forceinline__ void store_link(const DirView<T> &d, int idx, const M3<T> &m)
{
    using V2 = typename vec2<T>::type;
    const unsigned off = (unsigned)idx * (unsigned)sizeof(V2);
#pragma unroll
    for (int k = 0; k < 9; k++) {
        V2 z;
        z.x = m.e[k / 3][k % 3].r;
        z.y = m.e[k / 3][k % 3].i;
        buf_store(d, off, k, z);
    }
}

// Convenience forms for a wave-uniform mu.
template <class T>
__device__ __forceinline__ M3<T> load_link(const typename vec2<T>::type *f, int v, int mu, int idx)
{
    return load_link<T>(dir_view<T>(f, v, mu), idx);
}
template <class T>
__device__ __forceinline__ void store_link(typename vec2<T>::type *f, int v, int mu, int idx, const M3<T> &m)
{
    store_link<T>(dir_view<T>(f, v, mu), idx, m);
}

// Plain-pointer forms for kernels in which mu varies within a wavefront.
template <class T>
__device__ __forceinline__ void store_link_any(typename vec2<T>::type *__restrict__ f, int v, int mu, int idx,
                                               const M3<T> &m)
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

// Wilson staple (mu wave-uniform) A_mu(x) = sum_{nu != mu} [ U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
//                                        + U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu) ]
// (GLAC/actions/wilsongaugeaction.cpp:29-53).
template <class T>
__device__ __forceinline__ M3<T> staple(const typename vec2<T>::type *__restrict__ f, const Geom &g, const Coord &x,
                                        int mu)
{
    M3<T> a = zero3<T>();
    const Coord xpm = step(g, x, mu, +1);
    const int ix = index_of(g, x), ipm = index_of(g, xpm);
    const DirView<T> vmu = dir_view<T>(f, g.v, mu);
#pragma unroll 1
    for (int d = 1; d < 4; d++) {
        const int nu = (mu + d) & 3;
        const DirView<T> vnu = dir_view<T>(f, g.v, nu);
        const int ipn = index_of(g, step(g, x, nu, +1));
        const int imn = index_of(g, step(g, x, nu, -1));
        const int ipmmn = index_of(g, step(g, xpm, nu, -1));
        {
            const M3<T> u1 = load_link<T>(vnu, ipm);
            const M3<T> u2 = load_link<T>(vmu, ipn);
            const M3<T> u3 = load_link<T>(vnu, ix);
            mm<false, true, true>(a, mul_ad(u1, u2), u3);
        }
        {
            const M3<T> u1 = load_link<T>(vnu, ipmmn);
            const M3<T> u2 = load_link<T>(vmu, imn);
            const M3<T> u3 = load_link<T>(vnu, imn);
            mm<false, false, true>(a, mul_dd(u1, u2), u3);
        }
    }
    return a;
}

} // namespace glacx
