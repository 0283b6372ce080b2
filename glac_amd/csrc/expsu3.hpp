// expsu3.hpp -- SU(3) exponential of a traceless antihermitian matrix by
// Cayley-Hamilton (Morningstar & Peardon, PRD 69 054501), the method of
// GLAC's SU3Exp::exp (GLAC/math/exponentiation/su3exp.cpp:16-79).
//
// exp(Y) = exp(iH) = f0 1 + f1 H + f2 H^2 with H = -iY.  The invariants
// c0 = det H, c1 = tr H^2 / 2 and the coefficients f_j are always evaluated
// in FP64; only the final matrix assembly runs in the link precision T.
//
// For ||H||_F^2 = 2 c1 <= 1 (every flow step at eps <= 0.02 on physical
// fields) the f_j are summed from the power series of exp(iH), each power
// reduced by the Cayley-Hamilton relation of a traceless 3x3 matrix,
// H^3 = c1 H + c0 1:  H^n = a_n 1 + b_n H + d_n H^2 with
// (a, b, d)_{n+1} = (c0 d_n, a_n + c1 d_n, b_n),  (a, b, d)_0 = (1, 0, 0);
// the terms are bounded by ||H||^n / n!, so 12 terms reach FP64 rounding for
// 2 c1 <= 0.01 and 20 terms for 2 c1 <= 1.  Larger arguments use the closed
// form of Morningstar & Peardon with acos/sin/cos.  Two guards GLAC lacks:
// |c0| / c0max is clamped to 1 before acos, and Y -> 0 needs no division.
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

// The same closed form in FP32 (the learned proposal's step matrix with FP32 links).
__device__ __forceinline__ void exp_coefficients_f32(float c0, float c1, float fo[3][2])
{
    const float c0max = 2.0f * (c1 * (1.0f / 3.0f)) * sqrtf(c1 * (1.0f / 3.0f));
    float ratio = fabsf(c0) / c0max;
    ratio = ratio > 1.0f ? 1.0f : ratio;
    const float theta = acosf(ratio);
    float st, ct;
    sincosf(theta * (1.0f / 3.0f), &st, &ct);
    const float u = sqrtf(c1 * (1.0f / 3.0f)) * ct;
    const float w = sqrtf(c1) * st;
    const float uu = u * u, ww = w * w;
    float su, cu, sw, cw;
    sincosf(u, &su, &cu);
    sincosf(w, &sw, &cw);
    const float xi0 = fabsf(w) > 0.05f ? sw / w : 1.0f - ww * (1.0f / 6.0f) * (1.0f - ww * (1.0f / 20.0f) * (1.0f - ww * (1.0f / 42.0f)));
    const float c2u = cu * cu - su * su, s2u = 2.0f * su * cu;
    float h[3][2];
    {   // h0 = (u^2 - w^2) e^{2iu} + e^{-iu} [8u^2 cos w + 2iu(3u^2 + w^2) xi0]
        const float br = 8.0f * uu * cw, bi = 2.0f * u * (3.0f * uu + ww) * xi0;
        h[0][0] = (uu - ww) * c2u + (cu * br + su * bi);
        h[0][1] = (uu - ww) * s2u + (cu * bi - su * br);
    }
    {   // h1 = 2u e^{2iu} - e^{-iu} [2u cos w - i(3u^2 - w^2) xi0]
        const float br = 2.0f * u * cw, bi = -(3.0f * uu - ww) * xi0;
        h[1][0] = 2.0f * u * c2u - (cu * br + su * bi);
        h[1][1] = 2.0f * u * s2u - (cu * bi - su * br);
    }
    {   // h2 = e^{2iu} - e^{-iu} [cos w + 3iu xi0]
        const float br = cw, bi = 3.0f * u * xi0;
        h[2][0] = c2u - (cu * br + su * bi);
        h[2][1] = s2u - (cu * bi - su * br);
    }
    const float inv = 1.0f / (9.0f * uu - ww);
#pragma unroll
    for (int j = 0; j < 3; j++) {
        fo[j][0] = h[j][0] * inv;
        fo[j][1] = h[j][1] * inv;
    }
    if (c0 < 0.0f) {  // f_j(-c0) = (-1)^j conj(f_j(c0))
        fo[0][1] = -fo[0][1];
        fo[1][0] = -fo[1][0];
        fo[2][1] = -fo[2][1];
    }
}

// inverse factorials 1/n!, n = 0..20
__device__ constexpr double kInvFact[21] = {
    1.0, 1.0, 0.5, 1.6666666666666666e-01, 4.1666666666666664e-02, 8.3333333333333332e-03,
    1.3888888888888889e-03, 1.9841269841269841e-04, 2.4801587301587302e-05, 2.7557319223985893e-06,
    2.7557319223985888e-07, 2.5052108385441720e-08, 2.0876756987868100e-09, 1.6059043836821613e-10,
    1.1470745597729725e-11, 7.6471637318198164e-13, 4.7794773323873853e-14, 2.8114572543455206e-15,
    1.5619206968586225e-16, 8.2206352466243295e-18, 4.1103176233121648e-19};

template <int N> __device__ __forceinline__ ExpCoeffs exp_coefficients_series(double c0, double c1)
{
    double a = 1.0, b = 0.0, d = 0.0;
    ExpCoeffs r;
    r.f[0][0] = 1.0;
    r.f[0][1] = r.f[1][0] = r.f[1][1] = r.f[2][0] = r.f[2][1] = 0.0;
#pragma unroll
    for (int n = 1; n <= N; n++) {
        const double an = c0 * d, bn = fma(c1, d, a);
        d = b;
        a = an;
        b = bn;
        // i^n / n!: real for even n, imaginary for odd n, sign + + - -
        const double s = (n & 2) ? -kInvFact[n] : kInvFact[n];
        const int part = n & 1;
        r.f[0][part] = fma(s, a, r.f[0][part]);
        r.f[1][part] = fma(s, b, r.f[1][part]);
        r.f[2][part] = fma(s, d, r.f[2][part]);
    }
    return r;
}

// FP32 variant of the series for the learned proposal's step matrix with FP32
// links (N terms: truncation below ||H||^(N+1) / (N+1)!).
template <int N> __device__ __forceinline__ void exp_coefficients_series_f32(float c0, float c1, float f[3][2])
{
    float a = 1.f, b = 0.f, d = 0.f;
    f[0][0] = 1.f;
    f[0][1] = f[1][0] = f[1][1] = f[2][0] = f[2][1] = 0.f;
#pragma unroll
    for (int n = 1; n <= N; n++) {
        const float an = c0 * d, bn = fmaf(c1, d, a);
        d = b;
        a = an;
        b = bn;
        const float s = (float)((n & 2) ? -kInvFact[n] : kInvFact[n]);
        const int part = n & 1;
        f[0][part] = fmaf(s, a, f[0][part]);
        f[1][part] = fmaf(s, b, f[1][part]);
        f[2][part] = fmaf(s, d, f[2][part]);
    }
}

// X = f0 1 + f1 (-iY) + f2 (-Y^2) from the coefficients.
template <class T> __device__ __forceinline__ M3<T> exp_assemble(const M3<T> &Y, T f0r, T f0i, T f1r, T f1i, T f2r, T f2i)
{
    const M3<T> Y2 = square_antiherm(Y);
    M3<T> x;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            // f1 * (-i Y) = (f1r + i f1i)(Y.i - i Y.r) ;  -f2 Y2
            const T hr = Y.e[i][j].i, hi = -Y.e[i][j].r;
            x.e[i][j].r = fma(f1r, hr, fma(-f1i, hi, fma(-f2r, Y2.e[i][j].r, f2i * Y2.e[i][j].i)));
            x.e[i][j].i = fma(f1r, hi, fma(f1i, hr, fma(-f2r, Y2.e[i][j].i, -f2i * Y2.e[i][j].r)));
        }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        x.e[i][i].r += f0r;
        x.e[i][i].i += f0i;
    }
    return x;
}

template <class T> __device__ __forceinline__ M3<T> exp_su3(const su3alg<T> &y);

// Step matrix of the learned Metropolis proposal.  FP64 links use exp_su3.
// FP32 links evaluate everything in FP32: the series when ||H||_F^2 <= 4, the
// closed form otherwise.  The proposal only has to be an SU(3) matrix drawn
// symmetrically, which FP32 rounding of X(w) and X(-w) keeps to the same
// order as the RST generator's; links are reunitarised after every sweep.
template <class T> __device__ __forceinline__ M3<T> exp_su3_proposal(const su3alg<T> &y)
{
    if constexpr (sizeof(T) == 8) {
        return exp_su3(y);
    } else {
        const float a0 = y.a0, a1 = y.a1, a2 = -(a0 + a1);
        const float h01r = y.p[1], h01i = -y.p[0], h02r = y.p[3], h02i = -y.p[2], h12r = y.p[5], h12i = -y.p[4];
        const float n01 = fmaf(h01r, h01r, h01i * h01i), n02 = fmaf(h02r, h02r, h02i * h02i),
                    n12 = fmaf(h12r, h12r, h12i * h12i);
        const float c1 = 0.5f * (a0 * a0 + a1 * a1 + a2 * a2) + n01 + n02 + n12;
        const float tr = h01r * h12r - h01i * h12i, ti = h01r * h12i + h01i * h12r;
        const float c0 = a0 * a1 * a2 + 2.f * (tr * h02r + ti * h02i) - a0 * n12 - a1 * n02 - a2 * n01;
        // N terms leave about |h|^(N+1) / (N+1)! of the series, h the largest
        // eigenvalue of H in magnitude; with tr H = 0 and tr H^2 = 2 c1,
        // |h| <= sqrt(4 c1 / 3).  That is below FP32 rounding (6e-8) for
        // c1 <= 0.125 (N = 7), 0.5 (9), 1 (11), 2 (13); the wavefront takes
        // the largest order any of its series lanes needs (trained steps have
        // c1 ~ 0.4, so a wavefront of 64 usually holds a lane above 0.5).
        float f[3][2];
        if (c1 > 2.0f)
            exp_coefficients_f32(c0, c1, f);
        else if (__any(c1 > 1.0f))
            exp_coefficients_series_f32<13>(c0, c1, f);
        else if (__any(c1 > 0.5f))
            exp_coefficients_series_f32<11>(c0, c1, f);
        else if (__any(c1 > 0.125f))
            exp_coefficients_series_f32<9>(c0, c1, f);
        else
            exp_coefficients_series_f32<7>(c0, c1, f);
        return exp_assemble<float>(expand(y), f[0][0], f[0][1], f[1][0], f[1][1], f[2][0], f[2][1]);
    }
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
    // Series order: N terms leave a remainder below ||H||^(N+1) / (N+1)!,
    // ||H||_F^2 = 2 c1, under FP64 rounding for c1 <= 2e-6 (N = 5), 2e-4 (7),
    // 0.005 (9), 0.5 (20); the closed form beyond.  The order is the largest
    // any series lane of the wavefront needs, so the series is one path.
    ExpCoeffs f;
    const bool series = c1 <= 0.5;
    if (!series)
        f = exp_coefficients(c0, c1);
    else if (__any(c1 > 0.005))
        f = exp_coefficients_series<20>(c0, c1);
    else if (__any(c1 > 2e-4))
        f = exp_coefficients_series<9>(c0, c1);
    else if (__any(c1 > 2e-6))
        f = exp_coefficients_series<7>(c0, c1);
    else
        f = exp_coefficients_series<5>(c0, c1);
    // H = -iY, H^2 = -Y^2:  X = f0 1 + f1 (-iY) + f2 (-Y^2)
    return exp_assemble<T>(Y, T(f.f[0][0]), T(f.f[0][1]), T(f.f[1][0]), T(f.f[1][1]), T(f.f[2][0]), T(f.f[2][1]));
}

} // namespace glacx
