// k_field.hip -- field layout conversion, starts and the exported flow force.
//
//  pack/unpack: GLAC host layout (AoS FP64, mu-major or .bin site-major,
//               GLAC/io/fieldio.cpp:95-107, 226-235) <-> device SoA planes.
//  hot start:   from the counter-based stream (GLAC/system.cpp:200-214):
//               Haar SU(3), GLAC's "fully random" generateRandom (det = -1,
//               su3matrixgenerator.h:98-164) or its RST matrix near unity
//               (RSTHotStart, su3matrixgenerator.h:170-196).
//  action derivative: Z_mu for GLAC's getActionDerivative API
//               (GLAC/actions/wilsongaugeaction.cpp:55-86).
#include "kernels.hpp"
#include "proposal.hpp"
#include "rng.hpp"

namespace glacx {

template <class T, int R>
__global__ __launch_bounds__(kBlock) void pack_kernel(const double *__restrict__ stage,
                                                      typename vec2<T>::type *__restrict__ f, Geom g, int site_major)
{
    const int idx = blockIdx.x * kBlock + threadIdx.x;
    if (idx >= g.v) return;
    const int lex = lex_of(g, decode(g, idx));
#pragma unroll
    for (int mu = 0; mu < 4; mu++) {
        const double *src = site_major ? stage + ((size_t)lex * 4 + mu) * 18 : stage + ((size_t)mu * g.v + lex) * 18;
        M3<T> m;
#pragma unroll
        for (int k = 0; k < 9; k++) m.e[k / 3][k % 3] = {T(src[2 * k]), T(src[2 * k + 1])};
        store_link<T, R>(f, g, mu, idx, m);
    }
}

template <class T, int R>
__global__ __launch_bounds__(kBlock) void unpack_kernel(const typename vec2<T>::type *__restrict__ f,
                                                        double *__restrict__ stage, Geom g, int site_major)
{
    const int idx = blockIdx.x * kBlock + threadIdx.x;
    if (idx >= g.v) return;
    const int lex = lex_of(g, decode(g, idx));
#pragma unroll
    for (int mu = 0; mu < 4; mu++) {
        double *dst = site_major ? stage + ((size_t)lex * 4 + mu) * 18 : stage + ((size_t)mu * g.v + lex) * 18;
        const M3<T> m = load_link<T, R>(f, g, mu, idx);
#pragma unroll
        for (int k = 0; k < 9; k++) {
            dst[2 * k] = double(m.e[k / 3][k % 3].r);
            dst[2 * k + 1] = double(m.e[k / 3][k % 3].i);
        }
    }
}

// GLAC's generateRandom (GLAC/math/matrices/su3matrixgenerator.h:98-164) on
// the entries already in m: column 0 normalised, column 1 orthogonalised to it
// and normalised, column 2 = -conj(c0 x c1), which makes det = -1.
__device__ __forceinline__ void glac_random_columns(M3<double> &m)
{
    double n0 = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) n0 += m.e[i][0].r * m.e[i][0].r + m.e[i][0].i * m.e[i][0].i;
    n0 = sqrt(n0);
#pragma unroll
    for (int i = 0; i < 3; i++) m.e[i][0] = {m.e[i][0].r / n0, m.e[i][0].i / n0};
    double pr = 0.0, pi = 0.0;  // c0^dagger c1
#pragma unroll
    for (int i = 0; i < 3; i++) {
        pr += m.e[i][0].r * m.e[i][1].r + m.e[i][0].i * m.e[i][1].i;
        pi += m.e[i][0].r * m.e[i][1].i - m.e[i][0].i * m.e[i][1].r;
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double cr = m.e[i][0].r, ci = m.e[i][0].i;
        m.e[i][1].r -= cr * pr - ci * pi;
        m.e[i][1].i -= cr * pi + ci * pr;
    }
    double n1 = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) n1 += m.e[i][1].r * m.e[i][1].r + m.e[i][1].i * m.e[i][1].i;
    n1 = sqrt(n1);
#pragma unroll
    for (int i = 0; i < 3; i++) m.e[i][1] = {m.e[i][1].r / n1, m.e[i][1].i / n1};
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const int j = (i + 1) % 3, k = (i + 2) % 3;
        const cpx<double> x = cmul(m.e[j][0], m.e[k][1]), y = cmul(m.e[k][0], m.e[j][1]);
        m.e[i][2] = {-(x.r - y.r), x.i - y.i};
    }
}

// One thread per link, drawing from the link's stream (seed, lex, mu, hot
// start, 0):
//   HOT_HAAR:  9 complex Gaussians by Box-Muller (entry k from uniforms 2k,
//              2k+1), then FP64 row Gram-Schmidt with row 2 = conj(r0 x r1);
//   HOT_GLAC_RANDOM:  18 uniforms eps (-1, 1) (entry k: 2k real, 2k+1
//              imaginary), then glac_random_columns;
//   HOT_RST:   GLAC's RST matrix with step eps (the Metropolis proposal's draws).
template <class T, int R>
__global__ __launch_bounds__(kBlock) void hot_start_kernel(typename vec2<T>::type *__restrict__ f, Geom g, uint64_t seed,
                                                           int kind, double eps)
{
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= 4 * g.v) return;
    const int mu = t / g.v, idx = t - mu * g.v;
    const int lex = lex_of(g, decode(g, idx));
    Stream rs(seed, (uint32_t)lex, (uint32_t)mu, STREAM_HOT_START, 0);
    M3<double> m;
    if (kind == HOT_GLAC_RANDOM) {
#pragma unroll
        for (int k = 0; k < 9; k++) {
            const double re = eps * rs.symmetric();
            m.e[k / 3][k % 3] = {re, eps * rs.symmetric()};
        }
        glac_random_columns(m);
    } else if (kind == HOT_RST) {
        m = rst_proposal<double>(rs, sqrt(1.0 - eps * eps), eps);
    } else {
#pragma unroll
        for (int k = 0; k < 9; k++) {
            const double u1 = rs.open(), u2 = rs.open();
            const double r = sqrt(-2.0 * log(u1));
            double s, c;
            sincos(2.0 * M_PI * u2, &s, &c);
            m.e[k / 3][k % 3] = {r * c, r * s};
        }
        reunitarize(m);
    }
    store_link_any<T, R>(f, g, mu, idx, convert<T>(m));
}

template <class T, int R>
__global__ __launch_bounds__(kBlock) void cold_start_kernel(typename vec2<T>::type *__restrict__ f, Geom g)
{
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= 4 * g.v) return;
    const int mu = t / g.v, idx = t - mu * g.v;
    store_link_any<T, R>(f, g, mu, idx, unit3<T>());
}

template <class T, int R>
__global__ __launch_bounds__(kBlock) void action_derivative_kernel(const typename vec2<T>::type *__restrict__ f,
                                                                   double *__restrict__ out, Geom g, int mu)
{
    const int idx = blockIdx.x * kBlock + threadIdx.x;
    if (idx >= g.v) return;
    const Site x = site_of(g, idx);
    const M3<T> a = staple<T, R>(f, g, x, mu);
    const M3<T> u = load_link<T, R>(f, g, mu, idx);
    const M3<T> z = expand(force_from_omega(mul(u, a)));
    double *dst = out + (size_t)lex_of(g, x) * 18;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        dst[2 * k] = double(z.e[k / 3][k % 3].r);
        dst[2 * k + 1] = double(z.e[k / 3][k % 3].i);
    }
}

static inline unsigned blocks_for(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

template <class T, int R> void launch_pack(const double *stage, void *f, const Geom &g, int site_major, hipStream_t s)
{
    pack_kernel<T, R><<<blocks_for(g.v), kBlock, 0, s>>>(stage, (typename vec2<T>::type *)f, g, site_major);
}
template <class T, int R> void launch_unpack(const void *f, double *stage, const Geom &g, int site_major, hipStream_t s)
{
    unpack_kernel<T, R><<<blocks_for(g.v), kBlock, 0, s>>>((const typename vec2<T>::type *)f, stage, g, site_major);
}
template <class T, int R> void launch_hot_start(void *f, const Geom &g, uint64_t seed, int kind, double eps, hipStream_t s)
{
    hot_start_kernel<T, R><<<blocks_for(4L * g.v), kBlock, 0, s>>>((typename vec2<T>::type *)f, g, seed, kind, eps);
}
template <class T, int R> void launch_cold_start(void *f, const Geom &g, hipStream_t s)
{
    cold_start_kernel<T, R><<<blocks_for(4L * g.v), kBlock, 0, s>>>((typename vec2<T>::type *)f, g);
}
template <class T, int R> void launch_action_derivative(const void *f, double *stage, const Geom &g, int mu, hipStream_t s)
{
    action_derivative_kernel<T, R><<<blocks_for(g.v), kBlock, 0, s>>>((const typename vec2<T>::type *)f, stage, g, mu);
}

#define INSTANTIATE(T, R)                                                                                \
    template void launch_pack<T, R>(const double *, void *, const Geom &, int, hipStream_t);             \
    template void launch_unpack<T, R>(const void *, double *, const Geom &, int, hipStream_t);           \
    template void launch_hot_start<T, R>(void *, const Geom &, uint64_t, int, double, hipStream_t);      \
    template void launch_cold_start<T, R>(void *, const Geom &, hipStream_t);                            \
    template void launch_action_derivative<T, R>(const void *, double *, const Geom &, int, hipStream_t);
INSTANTIATE(float, 3)
INSTANTIATE(float, 2)
INSTANTIATE(double, 3)

} // namespace glacx
