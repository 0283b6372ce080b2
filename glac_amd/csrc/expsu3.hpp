// expsu3.hpp -- SU(3) exponential of a traceless antihermitian matrix by
// Cayley-Hamilton (Morningstar & Peardon, PRD 69 054501), the method of
// GLAC's SU3Exp::exp (GLAC/math/exponentiation/su3exp.cpp:16-79).
//
// exp(Y) = exp(iH) = f0 1 + f1 H + f2 H^2 with H = -iY.  The invariants
// c0 = det H, c1 = tr H^2 / 2 and the coefficients f_j (with their acos,
// sin and cos) are always evaluated in FP64; only the final matrix assembly
// runs in the link precision T.  Two guards GLAC lacks: |c0| / c0max is
// clamped to 1 before acos, and Y -> 0 (c1 < 1e-28, where GLAC divides 0/0)
// returns the second-order series, exact to rounding there.
#pragma once
#include <hip/hip_runtime.h>

#include "su3.hpp"

namespace glacx {

struct ExpCoeffs {
    double f[3][2];
};

// f_j of exp(iH) for Hermitian traceless H given c0 = det H, c1 = tr H^2 / 2 (c1 > 0).
__device__ __forceinline__ ExpCoeffs exp_coefficients(double c0, double c1)
{
    const double c0max = 2.0 * (c1 * (1.0 / 3.0)) * sqrt(c1 * (1.0 / 3.0));
    double ratio = fabs(c0) / c0max;
    ratio = ratio > 1.0 ? 1.0 : ratio;
    const double theta = acos(ratio);
    double st, ct;
    sincos(theta * (1.0 / 3.0), &st, &ct);
    const double u = sqrt(c1 * (1.0 / 3.0)) * ct;
    const double w = sqrt(c1) * st;
    const double uu = u * u, ww = w * w;
    double su, cu, sw, cw;
    sincos(u, &su, &cu);
    sincos(w, &sw, &cw);
    const double xi0 = fabs(w) > 0.05 ? sw / w : 1.0 - ww * (1.0 / 6.0) * (1.0 - ww * (1.0 / 20.0) * (1.0 - ww * (1.0 / 42.0)));
    const double c2u = cu * cu - su * su, s2u = 2.0 * su * cu;
    double h[3][2];
    {   // h0 = (u^2 - w^2) e^{2iu} + e^{-iu} [8u^2 cos w + 2iu(3u^2 + w^2) xi0]
        const double br = 8.0 * uu * cw, bi = 2.0 * u * (3.0 * uu + ww) * xi0;
        h[0][0] = (uu - ww) * c2u + (cu * br + su * bi);
        h[0][1] = (uu - ww) * s2u + (cu * bi - su * br);
    }
    {   // h1 = 2u e^{2iu} - e^{-iu} [2u cos w - i(3u^2 - w^2) xi0]
        const double br = 2.0 * u * cw, bi = -(3.0 * uu - ww) * xi0;
        h[1][0] = 2.0 * u * c2u - (cu * br + su * bi);
        h[1][1] = 2.0 * u * s2u - (cu * bi - su * br);
    }
    {   // h2 = e^{2iu} - e^{-iu} [cos w + 3iu xi0]
        const double br = cw, bi = 3.0 * u * xi0;
        h[2][0] = c2u - (cu * br + su * bi);
        h[2][1] = s2u - (cu * bi - su * br);
    }
    const double inv = 1.0 / (9.0 * uu - ww);
    ExpCoeffs r;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        r.f[j][0] = h[j][0] * inv;
        r.f[j][1] = h[j][1] * inv;
    }
    if (c0 < 0.0) {  // f_j(-c0) = (-1)^j conj(f_j(c0))
        r.f[0][1] = -r.f[0][1];
        r.f[1][0] = -r.f[1][0];
        r.f[2][1] = -r.f[2][1];
    }
    return r;
}

template <class T> __device__ __forceinline__ M3<T> exp_su3(const su3alg<T> &y)
{
    // H = -iY: H_jj = a_j, H_jk = -i Y_jk (j < k)
    const double a0 = y.a0, a1 = y.a1, a2 = -(a0 + a1);
    const double h01r = y.p[1], h01i = -(double)y.p[0];
    const double h02r = y.p[3], h02i = -(double)y.p[2];
    const double h12r = y.p[5], h12i = -(double)y.p[4];
    const double n01 = h01r * h01r + h01i * h01i;
    const double n02 = h02r * h02r + h02i * h02i;
    const double n12 = h12r * h12r + h12i * h12i;
    const double c1 = 0.5 * (a0 * a0 + a1 * a1 + a2 * a2) + n01 + n02 + n12;
    // det H = a0 a1 a2 + 2 Re(h01 h12 conj(h02)) - a0 |h12|^2 - a1 |h02|^2 - a2 |h01|^2
    const double tr = h01r * h12r - h01i * h12i, ti = h01r * h12i + h01i * h12r;
    const double c0 = a0 * a1 * a2 + 2.0 * (tr * h02r + ti * h02i) - a0 * n12 - a1 * n02 - a2 * n01;

    const M3<T> Y = expand(y);
    if (!(c1 > 1e-28)) {
        M3<T> x = mul(Y, Y);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                x.e[i][j].r = T(0.5) * x.e[i][j].r + Y.e[i][j].r;
                x.e[i][j].i = T(0.5) * x.e[i][j].i + Y.e[i][j].i;
            }
        x.e[0][0].r += T(1);
        x.e[1][1].r += T(1);
        x.e[2][2].r += T(1);
        return x;
    }
    const ExpCoeffs f = exp_coefficients(c0, c1);
    // H = -iY, H^2 = -Y^2:  X = f0 1 + f1 (-iY) + f2 (-Y^2)
    const M3<T> Y2 = mul(Y, Y);
    const T f1r = T(f.f[1][0]), f1i = T(f.f[1][1]), f2r = T(f.f[2][0]), f2i = T(f.f[2][1]);
    M3<T> x;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            // f1 * (-i Y) = (f1r + i f1i)(Y.i - i Y.r) ;  -f2 Y2
            const T hr = Y.e[i][j].i, hi = -Y.e[i][j].r;
            x.e[i][j].r = f1r * hr - f1i * hi - (f2r * Y2.e[i][j].r - f2i * Y2.e[i][j].i);
            x.e[i][j].i = f1r * hi + f1i * hr - (f2r * Y2.e[i][j].i + f2i * Y2.e[i][j].r);
        }
    x.e[0][0].r += T(f.f[0][0]);
    x.e[0][0].i += T(f.f[0][1]);
    x.e[1][1].r += T(f.f[0][0]);
    x.e[1][1].i += T(f.f[0][1]);
    x.e[2][2].r += T(f.f[0][0]);
    x.e[2][2].i += T(f.f[0][1]);
    return x;
}

} // namespace glacx
