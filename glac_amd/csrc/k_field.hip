// k_field.hip -- field layout conversion, starts and the exported flow force.
//
//  pack/unpack: GLAC host layout (AoS FP64, mu-major or .bin site-major,
//               GLAC/io/fieldio.cpp:95-107, 226-235) <-> device SoA planes.
//  hot start:   Haar SU(3) from the counter-based stream (replaces
//               GLAC/system.cpp:200-214; GLAC's own generator makes det = -1
//               matrices, su3matrixgenerator.h:156-161).
//  action derivative: Z_mu for GLAC's getActionDerivative API
//               (GLAC/actions/wilsongaugeaction.cpp:55-86).
#include "kernels.hpp"
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

// One thread per link: 9 complex Gaussians by Box-Muller (entry k from stream
// uniforms 2k, 2k+1), then FP64 row Gram-Schmidt with row 2 = conj(r0 x r1).
template <class T, int R>
__global__ __launch_bounds__(kBlock) void hot_start_kernel(typename vec2<T>::type *__restrict__ f, Geom g, uint64_t seed)
{
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= 4 * g.v) return;
    const int mu = t / g.v, idx = t - mu * g.v;
    const int lex = lex_of(g, decode(g, idx));
    Stream rs(seed, (uint32_t)lex, (uint32_t)mu, STREAM_HOT_START, 0);
    M3<double> m;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const double u1 = rs.open(), u2 = rs.open();
        const double r = sqrt(-2.0 * log(u1));
        double s, c;
        sincos(2.0 * M_PI * u2, &s, &c);
        m.e[k / 3][k % 3] = {r * c, r * s};
    }
    reunitarize(m);
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
    const Coord x = decode(g, idx);
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
template <class T, int R> void launch_hot_start(void *f, const Geom &g, uint64_t seed, hipStream_t s)
{
    hot_start_kernel<T, R><<<blocks_for(4L * g.v), kBlock, 0, s>>>((typename vec2<T>::type *)f, g, seed);
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
    template void launch_hot_start<T, R>(void *, const Geom &, uint64_t, hipStream_t);                   \
    template void launch_cold_start<T, R>(void *, const Geom &, hipStream_t);                            \
    template void launch_action_derivative<T, R>(const void *, double *, const Geom &, int, hipStream_t);
INSTANTIATE(float, 3)
INSTANTIATE(float, 2)
INSTANTIATE(double, 3)

} // namespace glacx
