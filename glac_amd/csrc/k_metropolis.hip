// k_metropolis.hip -- one checkerboard Metropolis sub-sweep (mu, parity).
//
// Replaces System::update / updateLink (GLAC/system.cpp:321-361): GLAC walks
// the links sequentially; here every link U_mu(x) with x of the given parity
// is updated at once.  Its staple only involves U_mu at the other parity and
// U_nu, nu != mu, none of which change during the sub-sweep, so the links are
// independent and are updated in place.
//
// Per link: the staple A (WilsonGaugeAction::computeStaple,
// GLAC/actions/wilsongaugeaction.cpp:29-53), then n_hit proposals U' = X U
// with X from GLAC's RST construction (GLAC/math/matrices/
// su3matrixgenerator.h:170-321): three SU(2) matrices x0 1 + i x.sigma in
// blocks (0,1), (0,2), (1,2), x0 = sign(r0) sqrt(1 - eps^2),
// x = eps r / |r|, r ~ U(-1/2, 1/2)^4, product R S T or its inverse by a coin.
// dS = -(beta/3) Re tr[(U' - U) A] (wilsongaugeaction.cpp:9, 24-27) is
// evaluated as -(beta/3) (Re tr(X W) - Re tr W) with W = U A kept up to date;
// accept iff exp(-dS) > u (GLAC/system.cpp:347).  After the hits the link is
// optionally reunitarised in FP64.  Uniforms come from the link's own
// counter-based stream (rng.hpp), 13 centred ones for the proposal and one
// open one for the test per hit, in that order.
#include "kernels.hpp"
#include "rng.hpp"

namespace glacx {

// SU(2) near unity from 4 centred uniforms; returns (a, b) with
// matrix [[a, b], [-conj(b), conj(a)]], a = x0 + i x3, b = x2 + i x1.
template <class T>
__device__ __forceinline__ void su2_near_unity(const T r0, const T r1, const T r2, const T r3, T x0abs, T eps,
                                               cpx<T> &a, cpx<T> &b)
{
    T x0 = r0 < T(0) ? -x0abs : x0abs;
    const T sc = eps * rsqrt(fma(r1, r1, fma(r2, r2, r3 * r3)));  // eps / |r|
    T x1 = r1 * sc, x2 = r2 * sc, x3 = r3 * sc;
    // |x| = 1 up to rounding (x0^2 = 1 - eps^2); renormalise as GLAC does
    const T inv = rsqrt(fma(x0, x0, fma(x1, x1, fma(x2, x2, x3 * x3))));
    x0 *= inv;
    x1 *= inv;
    x2 *= inv;
    x3 *= inv;
    a = {x0, x3};
    b = {x2, x1};
}

// X = R S T with r in block (0,1), s in block (0,2), t in block (1,2).
template <class T>
__device__ __forceinline__ M3<T> rst_product(cpx<T> ra, cpx<T> rb, cpx<T> sa, cpx<T> sb, cpx<T> ta, cpx<T> tb)
{
    // 2x2 entries: m00 = a, m01 = b, m10 = -conj(b), m11 = conj(a)
    const cpx<T> r00 = ra, r01 = rb, r10 = {-rb.r, rb.i}, r11 = {ra.r, -ra.i};
    const cpx<T> s00 = sa, s01 = sb, s10 = {-sb.r, sb.i}, s11 = {sa.r, -sa.i};
    const cpx<T> t00 = ta, t01 = tb, t10 = {-tb.r, tb.i}, t11 = {ta.r, -ta.i};
    const cpx<T> r00s01 = cmul(r00, s01), r10s01 = cmul(r10, s01);
    M3<T> x;
    x.e[0][0] = cmul(r00, s00);
    {
        const cpx<T> u = cmul(r01, t00), v = cmul(r00s01, t10);
        x.e[0][1] = {u.r + v.r, u.i + v.i};
    }
    {
        const cpx<T> u = cmul(r01, t01), v = cmul(r00s01, t11);
        x.e[0][2] = {u.r + v.r, u.i + v.i};
    }
    x.e[1][0] = cmul(r10, s00);
    {
        const cpx<T> u = cmul(r11, t00), v = cmul(r10s01, t10);
        x.e[1][1] = {u.r + v.r, u.i + v.i};
    }
    {
        const cpx<T> u = cmul(r11, t01), v = cmul(r10s01, t11);
        x.e[1][2] = {u.r + v.r, u.i + v.i};
    }
    x.e[2][0] = s10;
    x.e[2][1] = cmul(s11, t10);
    x.e[2][2] = cmul(s11, t11);
    return x;
}

// exp(-dS): FP64 libm exp for FP64 links; the hardware base-2 exponential for
// FP32 links, whose dS already carries FP32 rounding.
__device__ __forceinline__ double accept_prob(double ds) { return exp(-ds); }
__device__ __forceinline__ double accept_prob(float ds) { return (double)__expf(-ds); }

template <class T> __device__ __forceinline__ M3<T> dagger(const M3<T> &m)
{
    M3<T> d;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) d.e[i][j] = {m.e[j][i].r, -m.e[j][i].i};
    return d;
}

template <class T>
__global__ __launch_bounds__(kBlock) void metropolis_kernel(typename vec2<T>::type *__restrict__ f, Geom g, int mu,
                                                            int parity, MetroParams prm,
                                                            unsigned long long *__restrict__ count)
{
    const int h = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    double nacc[1] = {0.0};
    if (h < g.vh) {
        const int idx = parity_index(g, parity, h);
        const Coord x = decode(g, idx);
        const M3<T> a = staple<T>(f, g, x, mu);
        M3<T> u = load_link<T>(f, g.v, mu, idx);
        Stream rs(prm.seed, (uint32_t)lex_of(g, x), (uint32_t)mu, STREAM_METROPOLIS, prm.sweep);
        const T eps = T(prm.eps);
        const T x0abs = sqrt(T(1) - eps * eps);
        const double bf = prm.beta / 3.0;
        M3<T> w = mul(u, a);
        T trw = retrace(w);
        int accepted = 0;
#pragma unroll 1
        for (int hit = 0; hit < prm.n_hit; hit++) {
            cpx<T> ra, rb, sa, sb, ta, tb;
            {
                const T q0 = rs.centered<T>(), q1 = rs.centered<T>(), q2 = rs.centered<T>(), q3 = rs.centered<T>();
                su2_near_unity(q0, q1, q2, q3, x0abs, eps, ra, rb);
            }
            {
                const T q0 = rs.centered<T>(), q1 = rs.centered<T>(), q2 = rs.centered<T>(), q3 = rs.centered<T>();
                su2_near_unity(q0, q1, q2, q3, x0abs, eps, sa, sb);
            }
            {
                const T q0 = rs.centered<T>(), q1 = rs.centered<T>(), q2 = rs.centered<T>(), q3 = rs.centered<T>();
                su2_near_unity(q0, q1, q2, q3, x0abs, eps, ta, tb);
            }
            const T coin = rs.centered<T>();
            M3<T> xm = rst_product(ra, rb, sa, sb, ta, tb);
            if (coin < T(0)) xm = dagger(xm);
            const T ds = T(-bf) * (retrace_mul(xm, w) - trw);
            if (accept_prob(ds) > rs.open()) {
                u = mul(xm, u);
                w = mul(xm, w);
                trw = retrace(w);
                accepted++;
            }
        }
        if (prm.reunit) {
            M3<double> d = convert<double>(u);
            reunitarize(d);
            u = convert<T>(d);
        }
        store_link<T>(f, g.v, mu, idx, u);
        nacc[0] = (double)accepted;
    }
    block_sum<1>(nacc);
    if (threadIdx.x == 0) atomicAdd(count, (unsigned long long)nacc[0]);
}

template <class T>
void launch_metropolis(void *f, const Geom &g, int mu, int parity, const MetroParams &p, unsigned long long *count,
                       hipStream_t s)
{
    const unsigned blocks = (unsigned)((g.vh + kBlock - 1) / kBlock);
    metropolis_kernel<T><<<blocks, kBlock, 0, s>>>((typename vec2<T>::type *)f, g, mu, parity, p, count);
}

template void launch_metropolis<float>(void *, const Geom &, int, int, const MetroParams &, unsigned long long *,
                                       hipStream_t);
template void launch_metropolis<double>(void *, const Geom &, int, int, const MetroParams &, unsigned long long *,
                                        hipStream_t);

} // namespace glacx
