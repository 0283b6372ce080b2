// proposal.hpp -- GLAC's fixed-eps RST proposal and the accept test, shared by
// the plain and the learned Metropolis kernels.
//
// RST (GLAC/math/matrices/su3matrixgenerator.h:170-321): three SU(2)
// matrices x0 1 + i x.sigma, x0 = sign(r0) sqrt(1 - eps^2), x = eps r / |r|,
// r ~ U(-1/2, 1/2)^4, embedded in blocks (0,1), (0,2), (1,2); X = R S T, or
// its inverse when a 13th uniform is negative (draw order r, s, t, coin as
// GLAC).  Accept iff exp(-dS) > u (GLAC/system.cpp:347).
#pragma once
#include "lattice.hpp"
#include "rng.hpp"

namespace glacx {

// 1/sqrt for the proposal's normalisations: the hardware estimate (about
// 1 ulp) for FP32 links, whose products round at that level anyway; the
// correctly rounded library form for FP64 links.  (HIP's rsqrtf evaluates
// through FP64 to round correctly, about ten instructions per call.)
__device__ __forceinline__ float proposal_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double proposal_rsqrt(double x) { return rsqrt(x); }

// SU(2) near unity from 4 centred uniforms; returns (a, b) with
// matrix [[a, b], [-conj(b), conj(a)]], a = x0 + i x3, b = x2 + i x1.
template <class T>
__device__ __forceinline__ void su2_near_unity(const T r0, const T r1, const T r2, const T r3, T x0abs, T eps,
                                               cpx<T> &a, cpx<T> &b)
{
    T x0 = r0 < T(0) ? -x0abs : x0abs;
    const T sc = eps * proposal_rsqrt(fma(r1, r1, fma(r2, r2, r3 * r3)));  // eps / |r|
    T x1 = r1 * sc, x2 = r2 * sc, x3 = r3 * sc;
    // |x| = 1 up to rounding (x0^2 = 1 - eps^2).  FP64 renormalises as GLAC
    // does; in FP32 a renormalisation through the 1-ulp estimate would leave
    // |x| - 1 at the same ~1e-7 it finds, so it is skipped.
    if constexpr (sizeof(T) == 8) {
        const T inv = proposal_rsqrt(fma(x0, x0, fma(x1, x1, fma(x2, x2, x3 * x3))));
        x0 *= inv;
        x1 *= inv;
        x2 *= inv;
        x3 *= inv;
    }
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

// The proposed link U' = X U.  With two stored rows (SU(3) links) only rows 0
// and 1 are multiplied and row 2 is rebuilt as conj(row0 x row1).
template <class T, int R> __device__ __forceinline__ M3<T> propose_link(const M3<T> &xm, const M3<T> &u)
{
    if constexpr (R == 2) {
        M3<T> v;
        mm<false, false, false, 2>(v, xm, u);
        rebuild_third_row(v);
        return v;
    } else {
        return mul(xm, u);
    }
}

// GLAC's 13 uniforms in its order (r0..r3, s0..s3, t0..t3, coin) are the 16-bit
// halves of seven stream outputs, low half first (the 14th half is unused):
// seven 32-bit draws per proposal instead of thirteen.
template <class T> __device__ __forceinline__ M3<T> rst_proposal(Stream &rs, T x0abs, T eps)
{
    T q[14];
#pragma unroll
    for (int k = 0; k < 7; k++) rs.centered_pair<T>(q[2 * k], q[2 * k + 1]);
    cpx<T> ra, rb, sa, sb, ta, tb;
    su2_near_unity(q[0], q[1], q[2], q[3], x0abs, eps, ra, rb);
    su2_near_unity(q[4], q[5], q[6], q[7], x0abs, eps, sa, sb);
    su2_near_unity(q[8], q[9], q[10], q[11], x0abs, eps, ta, tb);
    const M3<T> xm = rst_product(ra, rb, sa, sb, ta, tb);
    return q[12] < T(0) ? dagger(xm) : xm;
}

} // namespace glacx
