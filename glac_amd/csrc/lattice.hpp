// lattice.hpp -- device layout of the gauge field and periodic neighbour
// arithmetic.
//
// Sites are stored in slabs: the z and t extents are cut into blocks of
// bz, bt = 4 sites (2 when the extent is not a multiple of 4; every extent is
// even), a block being bz bt full x-y planes.  Blocks are ordered z-block
// fastest; each block stores its even sites, then its odd sites, each half in
// the order (lt, lz, y, x / 2):
//     idx = (blk * 2 + p) * BH + ((lt * bz + lz) * Ny + y) * (Nx / 2) + (x >> 1),
//     blk = z / bz + (Nz / bz) (t / bt),  lz = z % bz,  lt = t % bt,
//     BH = bt bz Nx Ny / 2,  p = (x + y + z + t) & 1.
// Since Nx is even, x >> 1 is unique per parity within an x-row.
//
// Why this order: a wavefront of 64 lanes reads 64 consecutive entries of one
// plane.  For every neighbour direction the lanes' neighbours are then one
// or two contiguous runs (a step in y, z or t is a translation of the whole
// run; a step in x moves within the same rows), so each load touches close
// to the minimum number of cache lines: the texture address/data path, not
// HBM, bounds these kernels when a load is scattered over many lines.  The
// blocking in z and t keeps the t-neighbour of most sites (3 of 4) within
// bz Nx Ny / 2 entries, a few hundred KB of links at 32^4, inside an XCD's L2,
// where plain rows would put it a whole 3D slice away.
//
// A field is four direction regions of 9 V complex entries (mu-major).  In
// FP64 each region is 9 planes, entry (i, j) of U_mu at idx at
// plane[(mu * 9 + 3 i + j) * V + idx]; FP32 pairs entries into float4 planes
// (see load_rows01 below).
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

// Division by a runtime divisor 1 <= d < 2^30 for dividends 0 <= n < 2^30:
// q = (n * m) >> s with s = 30 + ceil(log2 d) and m = floor(2^s / d) + 1.
// With m = 2^s / d + e, 0 < e <= 1, n m / 2^s = n / d + n e / 2^s and
// n e / 2^s < 2^30 / 2^s <= 1 / d, which never reaches the next integer
// because the fractional part of n / d is at most (d - 1) / d.  n m < 2^62.
struct FastDiv {
    uint64_t m;
    int s;
    int d;
};

__host__ __device__ inline FastDiv make_fastdiv(int d)
{
    FastDiv f;
    int c = 0;
    while ((1 << c) < d) c++;
    f.d = d;
    f.s = 30 + c;
    f.m = ((uint64_t)1 << f.s) / (uint64_t)d + 1;
    return f;
}

__host__ __device__ inline int fdiv(int n, const FastDiv &f) { return (int)(((uint64_t)n * f.m) >> f.s); }

struct Geom {
    int n[4];       // Nx, Ny, Nz, Nt
    int nxh;        // Nx / 2
    int lbz, lbt;   // log2 of the block extents bz, bt
    int nzb;        // z-blocks, Nz / bz
    int run;        // entries of one (block, parity, lt): bz Nx Ny / 2
    int bh;         // entries of one (block, parity): bt run
    int v;          // V
    int vh;         // V / 2
    int vs;         // sites per time slice, Nx Ny Nz
    int cpr;        // 32-site chunks per run, ceil(run / 32)
    FastDiv d_nxh, d_ny, d_nzb, d_run, d_bh, d_2bh, d_cpr;
    // storage offsets of a step in y, z, t (index mu - 1), backwards [0] or
    // forwards [1], apart from the parity flip: within the block [0], into
    // the next block [1], across the periodic boundary [2] (site_step)
    int so[3][2][3];
};

// log2 of the preferred z and t block extents (a build-time tuning knob);
// an extent not divisible by the preferred block falls back to blocks of 2.
#ifndef GLACX_BLOCK_LOG_Z
#define GLACX_BLOCK_LOG_Z 2
#endif
#ifndef GLACX_BLOCK_LOG_T
#define GLACX_BLOCK_LOG_T 2
#endif

__host__ inline Geom make_geom(const int dims[4])
{
    Geom g;
    for (int k = 0; k < 4; k++) g.n[k] = dims[k];
    g.nxh = dims[0] / 2;
    g.lbz = (dims[2] % (1 << GLACX_BLOCK_LOG_Z) == 0) ? GLACX_BLOCK_LOG_Z : 1;
    g.lbt = (dims[3] % (1 << GLACX_BLOCK_LOG_T) == 0) ? GLACX_BLOCK_LOG_T : 1;
    g.nzb = dims[2] >> g.lbz;
    g.run = (1 << g.lbz) * dims[1] * g.nxh;
    g.bh = g.run << g.lbt;
    g.v = dims[0] * dims[1] * dims[2] * dims[3];
    g.vh = g.v / 2;
    g.vs = dims[0] * dims[1] * dims[2];
    g.cpr = (g.run + 31) / 32;
    g.d_cpr = make_fastdiv(g.cpr);
    g.d_nxh = make_fastdiv(g.nxh);
    g.d_ny = make_fastdiv(dims[1]);
    g.d_nzb = make_fastdiv(g.nzb);
    g.d_run = make_fastdiv(g.run);
    g.d_bh = make_fastdiv(g.bh);
    g.d_2bh = make_fastdiv(2 * g.bh);
    const int bz = 1 << g.lbz, bt = 1 << g.lbt, plane = dims[1] * g.nxh, ntb = dims[3] >> g.lbt;
    for (int f = 0; f < 2; f++) {
        const int s = f ? 1 : -1;
        // y: one row; across the boundary the other end of the plane
        g.so[0][f][0] = s * g.nxh;
        g.so[0][f][1] = g.so[0][f][2] = -s * (dims[1] - 1) * g.nxh;
        // z: one plane; into the next z-block (2 BH further) at its other end
        g.so[1][f][0] = s * plane;
        g.so[1][f][1] = -s * (bz - 1) * plane + s * 2 * g.bh;
        g.so[1][f][2] = -s * (bz - 1) * plane - s * 2 * g.bh * (g.nzb - 1);
        // t: one run of bz planes; into the next t-block (nzb z-blocks further)
        g.so[2][f][0] = s * g.run;
        g.so[2][f][1] = -s * (bt - 1) * g.run + s * 2 * g.bh * g.nzb;
        g.so[2][f][2] = -s * (bt - 1) * g.run - s * 2 * g.bh * g.nzb * (ntb - 1);
    }
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
    const int blk = fdiv(idx, g.d_2bh);
    int j = idx - blk * 2 * g.bh;
    const int p = j >= g.bh;
    j -= p * g.bh;
    const int q1 = fdiv(j, g.d_nxh);
    const int hx = j - q1 * g.nxh;
    const int q2 = fdiv(q1, g.d_ny);
    const int y = q1 - q2 * g.n[1];
    const int lz = q2 & ((1 << g.lbz) - 1);
    const int lt = q2 >> g.lbz;
    const int bt = fdiv(blk, g.d_nzb);
    const int bz = blk - bt * g.nzb;
    Coord r;
    r.c[1] = y;
    r.c[2] = (bz << g.lbz) + lz;
    r.c[3] = (bt << g.lbt) + lt;
    r.c[0] = 2 * hx + ((y + r.c[2] + r.c[3] + p) & 1);
    return r;
}

__host__ __device__ inline int lex_of(const Geom &g, const Coord &r)
{
    return r.c[0] + g.n[0] * (r.c[1] + g.n[1] * (r.c[2] + g.n[2] * r.c[3]));
}

__host__ __device__ inline int index_of(const Geom &g, const Coord &r)
{
    const int p = (r.c[0] + r.c[1] + r.c[2] + r.c[3]) & 1;
    const int blk = (r.c[2] >> g.lbz) + g.nzb * (r.c[3] >> g.lbt);
    const int l = ((r.c[3] & ((1 << g.lbt) - 1)) << g.lbz) + (r.c[2] & ((1 << g.lbz) - 1));
    return (blk * 2 + p) * g.bh + (l * g.n[1] + r.c[1]) * g.nxh + (r.c[0] >> 1);
}

// storage idx of the h-th site of parity p (h < V / 2): blocks in order, each
// contributing its BH sites of that parity.
__host__ __device__ inline int parity_index(const Geom &g, int p, int h)
{
    const int blk = fdiv(h, g.d_bh);
    return h + (blk + p) * g.bh;
}

// storage idx of the k-th site (k < run) of parity p in the run (blk, lt):
// all its sites share t = (blk / nzb) bt + lt.
__host__ __device__ inline int run_index(const Geom &g, int blk, int p, int lt, int k)
{
    return (blk * 2 + p) * g.bh + lt * g.run + k;
}

// storage idx of the r-th site of time slice t (r < Nx Ny Nz): the z-blocks
// of block row t / bt in order, each contributing its even, then its odd run
// of bz Nx Ny / 2 sites at lt = t % bt.
__host__ __device__ inline int slice_index(const Geom &g, int t, int r)
{
    const int q = fdiv(r, g.d_run);
    const int j = r - q * g.run;
    const int blk = (q >> 1) + g.nzb * (t >> g.lbt);
    return (blk * 2 + (q & 1)) * g.bh + (t & ((1 << g.lbt) - 1)) * g.run + j;
}

// ---- neighbour steps without multiplies ------------------------------------
// A site as its storage index plus the coordinates a step needs in the slab
// layout: x, y, the z and t offsets inside the block (lz, lt), the block's z
// and t indices (zb, tb) and the parity p.  A step flips the parity (the
// other half of the same block, +-BH) and moves within the block by a
// wave-uniform stride -- 1 per x pair, Nx/2 per y row, Ny Nx/2 per z plane,
// bz Ny Nx/2 per t slice -- or into the neighbouring block; every offset is a
// wave-uniform constant chosen per lane by a compare, so a step costs a few
// adds and selects where index_of(step(...)) costs three 32-bit multiplies
// (quarter-rate instructions).
struct Site {
    int idx, x, y, lz, zb, lt, tb, p;
};

__host__ __device__ inline Site site_of(const Geom &g, int idx)
{
    const int blk = fdiv(idx, g.d_2bh);
    int j = idx - blk * 2 * g.bh;
    const int p = j >= g.bh;
    j -= p * g.bh;
    const int q1 = fdiv(j, g.d_nxh);
    const int hx = j - q1 * g.nxh;
    const int q2 = fdiv(q1, g.d_ny);
    Site s;
    s.idx = idx;
    s.y = q1 - q2 * g.n[1];
    s.lz = q2 & ((1 << g.lbz) - 1);
    s.lt = q2 >> g.lbz;
    s.tb = fdiv(blk, g.d_nzb);
    s.zb = blk - s.tb * g.nzb;
    s.p = p;
    const int z = (s.zb << g.lbz) + s.lz, t = (s.tb << g.lbt) + s.lt;
    s.x = 2 * hx + ((s.y + z + t + p) & 1);
    return s;
}

// GLAC's lexicographic index x + Nx (y + Ny (z + Nz t)) of a site
__host__ __device__ inline int lex_of(const Geom &g, const Site &a)
{
    const int z = (a.zb << g.lbz) + a.lz, t = (a.tb << g.lbt) + a.lt;
    return a.x + g.n[0] * (a.y + g.n[1] * (z + g.n[2] * t));
}

// the site one step from `a` in direction mu (s = +1 or -1), periodic
__host__ __device__ inline Site site_step(const Geom &g, const Site &a, int mu, int s)
{
    Site b = a;
    b.p = a.p ^ 1;
    int off = a.p ? -g.bh : g.bh;
    const int f = s > 0;
    if (mu == 0) {
        int x = a.x + s;
        x = x < 0 ? g.n[0] - 1 : (x == g.n[0] ? 0 : x);
        b.x = x;
        off += (x >> 1) - (a.x >> 1);
    } else if (mu == 1) {
        const bool wrap = s > 0 ? a.y == g.n[1] - 1 : a.y == 0;
        b.y = wrap ? (s > 0 ? 0 : g.n[1] - 1) : a.y + s;
        off += wrap ? g.so[0][f][2] : g.so[0][f][0];
    } else if (mu == 2) {
        const int bz = 1 << g.lbz;
        const bool cross = s > 0 ? a.lz == bz - 1 : a.lz == 0;
        const bool wrap = s > 0 ? a.zb == g.nzb - 1 : a.zb == 0;
        b.lz = cross ? (s > 0 ? 0 : bz - 1) : a.lz + s;
        b.zb = cross ? (wrap ? (s > 0 ? 0 : g.nzb - 1) : a.zb + s) : a.zb;
        off += cross ? (wrap ? g.so[1][f][2] : g.so[1][f][1]) : g.so[1][f][0];
    } else {
        const int bt = 1 << g.lbt, ntb = g.n[3] >> g.lbt;
        const bool cross = s > 0 ? a.lt == bt - 1 : a.lt == 0;
        const bool wrap = s > 0 ? a.tb == ntb - 1 : a.tb == 0;
        b.lt = cross ? (s > 0 ? 0 : bt - 1) : a.lt + s;
        b.tb = cross ? (wrap ? (s > 0 ? 0 : ntb - 1) : a.tb + s) : a.tb;
        off += cross ? (wrap ? g.so[2][f][2] : g.so[2][f][1]) : g.so[2][f][0];
    }
    b.idx = a.idx + off;
    return b;
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
//
// Stored rows R (a template parameter of every kernel that touches links):
// R = 3 stores the whole matrix; R = 2 (the compressed FP32 mode) stores and
// reads only the first two rows and rebuilds the third as
// U_2j = conj(U_0k U_1l - U_0l U_1k), (j, k, l) cyclic, which holds for every
// SU(3) matrix (orthonormal rows, determinant 1).  A third of the link bytes
// is then never moved.
constexpr int kBufferFlags = 0x00020000;  // gfx9 raw buffer: 32-bit data format, no swizzle

template <class T> struct DirView {
    __amdgpu_buffer_rsrc_t r;
    int plane_bytes;
};

template <class T>
__device__ __forceinline__ DirView<T> dir_view(const typename vec2<T>::type *f, const Geom &g, int mu)
{
    using V2 = typename vec2<T>::type;
    DirView<T> d;
    const unsigned range = 9u * (unsigned)g.v * (unsigned)sizeof(V2);
    d.r = __builtin_amdgcn_make_buffer_rsrc((void *)(f + (size_t)mu * 9 * g.v), (short)0, (int)range, kBufferFlags);
    d.plane_bytes = g.v * (int)sizeof(V2);
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

// Third row of an SU(3) matrix from its first two.
template <class T> __host__ __device__ __forceinline__ void rebuild_third_row(M3<T> &m)
{
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int k = (j + 1) % 3, l = (j + 2) % 3;
        const cpx<T> a = m.e[0][k], b = m.e[1][l], c = m.e[0][l], d = m.e[1][k];
        const T re = fma(a.r, b.r, fma(-a.i, b.i, fma(-c.r, d.r, c.i * d.i)));
        const T im = fma(a.r, b.i, fma(a.i, b.r, fma(-c.r, d.i, -(c.i * d.r))));
        m.e[2][j] = {re, -im};
    }
}

// FP32 links pair their entries: plane pair q holds entries 2q and 2q + 1 of
// every link as one float4 (entry e = 3 i + j is U_ij), so the first two rows
// are three 16-byte loads; with three stored rows a fourth pair holds
// entries 6, 7 and entry 8 follows as a float2 plane (byte offset 64 V).  The
// direction's region is 72 V bytes either way; compressed links (R = 2) use
// its first 48 V.  FP64 links keep one double2 plane per entry.
__device__ __forceinline__ float4 buf_load_pair(const DirView<float> &d, unsigned off16, int q)
{
    return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(d.r, off16, q * 2 * d.plane_bytes, 0));
}
__device__ __forceinline__ void buf_store_pair(const DirView<float> &d, unsigned off16, int q, float4 z)
{
    using U4 = unsigned int __attribute__((ext_vector_type(4)));
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(U4, z), d.r, off16, q * 2 * d.plane_bytes, 0);
}

// First two rows only (row 2 left unset), for products that need no more of
// this operand.
template <class T, int R> __device__ __forceinline__ M3<T> load_rows01(const DirView<T> &d, int idx)
{
    M3<T> m;
    if constexpr (sizeof(T) == 4) {
        const unsigned off = (unsigned)idx * 16u;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const float4 z = buf_load_pair(d, off, q);
            m.e[(2 * q) / 3][(2 * q) % 3] = {z.x, z.y};
            m.e[(2 * q + 1) / 3][(2 * q + 1) % 3] = {z.z, z.w};
        }
    } else {
        using V2 = typename vec2<T>::type;
        const unsigned off = (unsigned)idx * (unsigned)sizeof(V2);
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const V2 z = buf_load(d, off, k);
            m.e[k / 3][k % 3] = {z.x, z.y};
        }
    }
    return m;
}

template <class T, int R> __device__ __forceinline__ M3<T> load_link(const DirView<T> &d, int idx)
{
    static_assert(R == 2 || R == 3, "links store 2 or 3 rows");
    static_assert(R == 3 || sizeof(T) == 4, "compressed links are FP32");
    M3<T> m = load_rows01<T, R>(d, idx);
    if constexpr (R == 2) {
        rebuild_third_row(m);
    } else if constexpr (sizeof(T) == 4) {
        const float4 z = buf_load_pair(d, (unsigned)idx * 16u, 3);
        m.e[2][0] = {z.x, z.y};
        m.e[2][1] = {z.z, z.w};
        const float2 w = buf_load(d, (unsigned)idx * 8u, 8);
        m.e[2][2] = {w.x, w.y};
    } else {
        using V2 = typename vec2<T>::type;
        const unsigned off = (unsigned)idx * (unsigned)sizeof(V2);
#pragma unroll
        for (int k = 6; k < 9; k++) {
            const V2 z = buf_load(d, off, k);
            m.e[2][k - 6] = {z.x, z.y};
        }
    }
    return m;
}

template <class T, int R> __device__ __forceinline__ void store_link(const DirView<T> &d, int idx, const M3<T> &m)
{
    if constexpr (sizeof(T) == 4) {
        const unsigned off = (unsigned)idx * 16u;
#pragma unroll
        for (int q = 0; q < (R == 2 ? 3 : 4); q++) {
            const cpx<T> a = m.e[(2 * q) / 3][(2 * q) % 3], b = m.e[(2 * q + 1) / 3][(2 * q + 1) % 3];
            buf_store_pair(d, off, q, make_float4(a.r, a.i, b.r, b.i));
        }
        if constexpr (R == 3) buf_store(d, (unsigned)idx * 8u, 8, make_float2(m.e[2][2].r, m.e[2][2].i));
    } else {
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
}

// Convenience forms for a wave-uniform mu.
template <class T, int R>
__device__ __forceinline__ M3<T> load_link(const typename vec2<T>::type *f, const Geom &g, int mu, int idx)
{
    return load_link<T, R>(dir_view<T>(f, g, mu), idx);
}
template <class T, int R>
__device__ __forceinline__ void store_link(typename vec2<T>::type *f, const Geom &g, int mu, int idx, const M3<T> &m)
{
    store_link<T, R>(dir_view<T>(f, g, mu), idx, m);
}

// Plain-pointer form for kernels in which mu varies within a wavefront.
template <class T, int R>
__device__ __forceinline__ void store_link_any(typename vec2<T>::type *__restrict__ f, const Geom &g, int mu, int idx,
                                               const M3<T> &m)
{
    typename vec2<T>::type *p = f + (size_t)mu * 9 * g.v;
    if constexpr (sizeof(T) == 4) {
        float4 *q4 = reinterpret_cast<float4 *>(p) + idx;
#pragma unroll
        for (int q = 0; q < (R == 2 ? 3 : 4); q++) {
            const cpx<T> a = m.e[(2 * q) / 3][(2 * q) % 3], b = m.e[(2 * q + 1) / 3][(2 * q + 1) % 3];
            q4[(size_t)q * g.v] = make_float4(a.r, a.i, b.r, b.i);
        }
        if constexpr (R == 3) p[(size_t)8 * g.v + idx] = make_float2(m.e[2][2].r, m.e[2][2].i);
    } else {
#pragma unroll
        for (int k = 0; k < 9; k++) {
            typename vec2<T>::type z;
            z.x = m.e[k / 3][k % 3].r;
            z.y = m.e[k / 3][k % 3].i;
            p[(size_t)k * g.v + idx] = z;
        }
    }
}

// Wilson staple (mu wave-uniform) A_mu(x) = sum_{nu != mu} [ U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
//                                        + U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu) ]
// (GLAC/actions/wilsongaugeaction.cpp:29-53).
template <class T, int R>
__device__ __forceinline__ M3<T> staple(const typename vec2<T>::type *__restrict__ f, const Geom &g, const Site &x,
                                        int mu)
{
    M3<T> a = zero3<T>();
    const Site xpm = site_step(g, x, mu, +1);
    const int ix = x.idx, ipm = xpm.idx;
    const DirView<T> vmu = dir_view<T>(f, g, mu);
#pragma unroll 1
    for (int d = 1; d < 4; d++) {
        const int nu = (mu + d) & 3;
        const DirView<T> vnu = dir_view<T>(f, g, nu);
        const int ipn = site_step(g, x, nu, +1).idx;
        const int imn = site_step(g, x, nu, -1).idx;
        const int ipmmn = site_step(g, xpm, nu, -1).idx;
        if constexpr (R == 2) {
            // SU(3) links (the compressed mode checks them): each term's
            // partial products keep two rows, the term's third is rebuilt
            M3<T> p, t;
            mm<false, true, false, 2>(p, load_rows01<T, R>(vnu, ipm), load_link<T, R>(vmu, ipn));
            mm<false, true, false, 2>(t, p, load_link<T, R>(vnu, ix));
            rebuild_third_row(t);
            add_to(a, t);
            mm<true, true, false, 2>(p, load_link<T, R>(vnu, ipmmn), load_link<T, R>(vmu, imn));
            mm<false, false, false, 2>(t, p, load_link<T, R>(vnu, imn));
            rebuild_third_row(t);
            add_to(a, t);
        } else {
            {
                const M3<T> u1 = load_link<T, R>(vnu, ipm);
                const M3<T> u2 = load_link<T, R>(vmu, ipn);
                const M3<T> u3 = load_link<T, R>(vnu, ix);
                mm<false, true, true>(a, mul_ad(u1, u2), u3);
            }
            {
                const M3<T> u1 = load_link<T, R>(vnu, ipmmn);
                const M3<T> u2 = load_link<T, R>(vmu, imn);
                const M3<T> u3 = load_link<T, R>(vnu, imn);
                mm<false, false, true>(a, mul_dd(u1, u2), u3);
            }
        }
    }
    return a;
}

} // namespace glacx
