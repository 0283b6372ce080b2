// su3.hpp -- register-resident 3x3 complex algebra for the gfx950 kernels.
//
// A link is held as 9 complex entries in registers, row-major.  All products
// are written out so the compiler schedules them as independent FMAs; the
// real type T is float (FP32 links) or double (FP64 links).
//
// Products follow GLAC's SU3 operators (GLAC/math/matrices/su3.h:184-221,
// 307-337) and traceRealMultiplication (GLAC/math/functions.h:45-53).
#pragma once
#include <hip/hip_runtime.h>

namespace glacx {

template <class T> struct cpx {
    T r, i;
};

template <class T> struct M3 {
    cpx<T> e[3][3];
};

template <class T> __device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b)
{
    return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}

// a * conj(b)
template <class T> __device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b)
{
    return {a.r * b.r + a.i * b.i, a.i * b.r - a.r * b.i};
}

// conj(a) * conj(b)
template <class T> __device__ __forceinline__ cpx<T> ccmulc(cpx<T> a, cpx<T> b)
{
    return {a.r * b.r - a.i * b.i, -(a.r * b.i + a.i * b.r)};
}

// conj(a) * b
template <class T> __device__ __forceinline__ cpx<T> cmulc_l(cpx<T> a, cpx<T> b)
{
    return {a.r * b.r + a.i * b.i, a.r * b.i - a.i * b.r};
}

template <class T> __device__ __forceinline__ M3<T> zero3()
{
    M3<T> m;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) m.e[i][j] = {T(0), T(0)};
    return m;
}

template <class T> __device__ __forceinline__ M3<T> unit3()
{
    M3<T> m = zero3<T>();
    m.e[0][0].r = m.e[1][1].r = m.e[2][2].r = T(1);
    return m;
}

// Generic 3x3 complex product written as FMA chains (one multiply and five
// fused multiply-adds per real entry, or six FMAs when accumulating).
// Element (i, k) of the left operand is a[i][k], or conj(a[k][i]) when DA is
// set; element (k, j) of the right operand is b[k][j], or conj(b[j][k]) when
// DB is set.  ACC adds the product to c instead of overwriting it.
// NR < 3 computes only the first NR rows of the product (the rest of c is
// left untouched): for a product of SU(3) matrices the third row follows from
// the first two (rebuild_third_row in lattice.hpp).
template <bool DA, bool DB, bool ACC, int NR = 3, class T>
__device__ __forceinline__ void mm(M3<T> &c, const M3<T> &a, const M3<T> &b)
{
    constexpr T sa = DA ? T(-1) : T(1);  // sign of the imaginary part of the left element
    constexpr T sb = DB ? T(-1) : T(1);
#pragma unroll
    for (int i = 0; i < NR; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            T re, im;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const cpx<T> x = DA ? a.e[k][i] : a.e[i][k];
                const cpx<T> y = DB ? b.e[j][k] : b.e[k][j];
                // (x.r + i sa x.i)(y.r + i sb y.i)
                if (k == 0 && !ACC) {
                    re = x.r * y.r;
                    im = (sb * x.r) * y.i;
                } else {
                    re = fma(x.r, y.r, k == 0 ? c.e[i][j].r : re);
                    im = fma(sb * x.r, y.i, k == 0 ? c.e[i][j].i : im);
                }
                re = fma(-(sa * sb) * x.i, y.i, re);
                im = fma(sa * x.i, y.r, im);
            }
            c.e[i][j] = {re, im};
        }
}

// C = A B
template <class T> __device__ __forceinline__ M3<T> mul(const M3<T> &a, const M3<T> &b)
{
    M3<T> c;
    mm<false, false, false>(c, a, b);
    return c;
}

// C = A B^dagger
template <class T> __device__ __forceinline__ M3<T> mul_ad(const M3<T> &a, const M3<T> &b)
{
    M3<T> c;
    mm<false, true, false>(c, a, b);
    return c;
}

// C = A^dagger B
template <class T> __device__ __forceinline__ M3<T> mul_da(const M3<T> &a, const M3<T> &b)
{
    M3<T> c;
    mm<true, false, false>(c, a, b);
    return c;
}

// C = A^dagger B^dagger = (B A)^dagger
template <class T> __device__ __forceinline__ M3<T> mul_dd(const M3<T> &a, const M3<T> &b)
{
    M3<T> c;
    mm<true, true, false>(c, a, b);
    return c;
}

template <class T> __device__ __forceinline__ void add_to(M3<T> &c, const M3<T> &a)
{
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            c.e[i][j].r += a.e[i][j].r;
            c.e[i][j].i += a.e[i][j].i;
        }
}

template <class T> __device__ __forceinline__ void negate(M3<T> &c)
{
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) c.e[i][j] = {-c.e[i][j].r, -c.e[i][j].i};
}

template <class T> __device__ __forceinline__ void sub_from(M3<T> &c, const M3<T> &a)
{
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            c.e[i][j].r -= a.e[i][j].r;
            c.e[i][j].i -= a.e[i][j].i;
        }
}

// Re tr(A B)
template <class T> __device__ __forceinline__ T retrace_mul(const M3<T> &a, const M3<T> &b)
{
    T s = a.e[0][0].r * b.e[0][0].r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            if (i != 0 || k != 0) s = fma(a.e[i][k].r, b.e[k][i].r, s);
            s = fma(-a.e[i][k].i, b.e[k][i].i, s);
        }
    return s;
}

// Y^2 for antihermitian Y (the result is hermitian: only the upper triangle
// is computed, the lower one is its conjugate).
template <class T> __device__ __forceinline__ M3<T> square_antiherm(const M3<T> &y)
{
    M3<T> c;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        // (Y^2)_ii = -sum_k |Y_ik|^2
        T d = -y.e[i][0].r * y.e[i][0].r;
        d = fma(-y.e[i][0].i, y.e[i][0].i, d);
        d = fma(-y.e[i][1].r, y.e[i][1].r, d);
        d = fma(-y.e[i][1].i, y.e[i][1].i, d);
        d = fma(-y.e[i][2].r, y.e[i][2].r, d);
        d = fma(-y.e[i][2].i, y.e[i][2].i, d);
        c.e[i][i] = {d, T(0)};
#pragma unroll
        for (int j = i + 1; j < 3; j++) {
            T re = y.e[i][0].r * y.e[0][j].r, im = y.e[i][0].r * y.e[0][j].i;
            re = fma(-y.e[i][0].i, y.e[0][j].i, re);
            im = fma(y.e[i][0].i, y.e[0][j].r, im);
#pragma unroll
            for (int k = 1; k < 3; k++) {
                re = fma(y.e[i][k].r, y.e[k][j].r, re);
                re = fma(-y.e[i][k].i, y.e[k][j].i, re);
                im = fma(y.e[i][k].r, y.e[k][j].i, im);
                im = fma(y.e[i][k].i, y.e[k][j].r, im);
            }
            c.e[i][j] = {re, im};
            c.e[j][i] = {re, -im};
        }
    }
    return c;
}

template <class T> __device__ __forceinline__ T retrace(const M3<T> &a)
{
    return a.e[0][0].r + a.e[1][1].r + a.e[2][2].r;
}

template <class To, class From> __device__ __forceinline__ M3<To> convert(const M3<From> &a)
{
    M3<To> c;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) c.e[i][j] = {To(a.e[i][j].r), To(a.e[i][j].i)};
    return c;
}

// T(M) = (M - M^dagger) - 1/3 tr(M - M^dagger), the traceless antihermitian
// part the clover observables take of C (GLAC/observables/
// mastersamplertopcxyz.cpp:331-340), of a product or a matrix, in the
// compact form {Im T00, Im T11, T01, T02, T12} (8 reals: T is traceless
// antihermitian, so T22 = -i (Im T00 + Im T11) and T_ji = -conj(T_ij)); the
// layout is that of su3alg.

// T(op(A) op(B)), op = dagger where the flag is set (the real diagonal of the
// product is not needed and not computed).
template <bool DA, bool DB = false, class T>
__device__ __forceinline__ void ta2_product(const M3<T> &a, const M3<T> &b, T f[8])
{
    M3<T> m;
    mm<DA, DB, false>(m, a, b);
    const T third = (m.e[0][0].i + m.e[1][1].i + m.e[2][2].i) * T(2.0 / 3.0);
    f[0] = T(2) * m.e[0][0].i - third;
    f[1] = T(2) * m.e[1][1].i - third;
    f[2] = m.e[0][1].r - m.e[1][0].r;
    f[3] = m.e[0][1].i + m.e[1][0].i;
    f[4] = m.e[0][2].r - m.e[2][0].r;
    f[5] = m.e[0][2].i + m.e[2][0].i;
    f[6] = m.e[1][2].r - m.e[2][1].r;
    f[7] = m.e[1][2].i + m.e[2][1].i;
}

// T(M) from the matrix itself.
template <class T> __device__ __forceinline__ void ta2_of(const M3<T> &m, T f[8])
{
    const T third = (m.e[0][0].i + m.e[1][1].i + m.e[2][2].i) * T(2.0 / 3.0);
    f[0] = T(2) * m.e[0][0].i - third;
    f[1] = T(2) * m.e[1][1].i - third;
    f[2] = m.e[0][1].r - m.e[1][0].r;
    f[3] = m.e[0][1].i + m.e[1][0].i;
    f[4] = m.e[0][2].r - m.e[2][0].r;
    f[5] = m.e[0][2].i + m.e[2][0].i;
    f[6] = m.e[1][2].r - m.e[2][1].r;
    f[7] = m.e[1][2].i + m.e[2][1].i;
}

// The full matrix of a compact T.
template <class T> __device__ __forceinline__ M3<T> ta2_matrix(const T f[8])
{
    M3<T> m;
    m.e[0][0] = {T(0), f[0]};
    m.e[1][1] = {T(0), f[1]};
    m.e[2][2] = {T(0), -(f[0] + f[1])};
    m.e[0][1] = {f[2], f[3]};
    m.e[1][0] = {-f[2], f[3]};
    m.e[0][2] = {f[4], f[5]};
    m.e[2][0] = {-f[4], f[5]};
    m.e[1][2] = {f[6], f[7]};
    m.e[2][1] = {-f[6], f[7]};
    return m;
}

// Re tr(F G) of two traceless antihermitian matrices in compact form.
template <class T> __device__ __forceinline__ T retrace_ta2(const T f[8], const T g[8])
{
    T s = f[0] * g[0];
    s = fma(f[1], g[1], s);
    s = fma(f[0] + f[1], g[0] + g[1], s);
    T o = f[2] * g[2];
#pragma unroll
    for (int k = 3; k < 8; k++) o = fma(f[k], g[k], o);
    return -s - T(2) * o;
}

// Compact su(3) element: traceless antihermitian Y with
//   Y_00 = i a0, Y_11 = i a1, Y_22 = -i (a0 + a1),
//   Y_01 = (p[0], p[1]), Y_02 = (p[2], p[3]), Y_12 = (p[4], p[5]),
//   Y_kj = -conj(Y_jk).
template <class T> struct su3alg {
    T a0, a1, p[6];
};

// Flow force Z = 1/2 [ (O^dag - O) - 1/3 Im tr(O^dag - O) i 1 ],  O = U A
// (GLAC/actions/wilsongaugeaction.cpp:77-85).
template <class T> __device__ __forceinline__ su3alg<T> force_from_omega(const M3<T> &o)
{
    su3alg<T> z;
    T tr = (o.e[0][0].i + o.e[1][1].i + o.e[2][2].i) * T(1.0 / 3.0);
    z.a0 = tr - o.e[0][0].i;
    z.a1 = tr - o.e[1][1].i;
    // Z_jk = 1/2 (conj(O_kj) - O_jk)
    z.p[0] = T(0.5) * (o.e[1][0].r - o.e[0][1].r);
    z.p[1] = T(-0.5) * (o.e[1][0].i + o.e[0][1].i);
    z.p[2] = T(0.5) * (o.e[2][0].r - o.e[0][2].r);
    z.p[3] = T(-0.5) * (o.e[2][0].i + o.e[0][2].i);
    z.p[4] = T(0.5) * (o.e[2][1].r - o.e[1][2].r);
    z.p[5] = T(-0.5) * (o.e[2][1].i + o.e[1][2].i);
    return z;
}

template <class T> __device__ __forceinline__ M3<T> expand(const su3alg<T> &y)
{
    M3<T> m;
    m.e[0][0] = {T(0), y.a0};
    m.e[1][1] = {T(0), y.a1};
    m.e[2][2] = {T(0), -(y.a0 + y.a1)};
    m.e[0][1] = {y.p[0], y.p[1]};
    m.e[0][2] = {y.p[2], y.p[3]};
    m.e[1][2] = {y.p[4], y.p[5]};
    m.e[1][0] = {-y.p[0], y.p[1]};
    m.e[2][0] = {-y.p[2], y.p[3]};
    m.e[2][1] = {-y.p[4], y.p[5]};
    return m;
}

// 1 / sqrt(x) for FP64 data headed for FP32 storage: the hardware's
// reciprocal square root estimate refined by two Newton steps (each squares
// the relative error), about a third of the instructions of the correctly
// rounded 1.0 / sqrt(x).
__device__ __forceinline__ double rsqrt_refined(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    y = y * fma(-0.5 * x * y, y, 1.5);
    return y * fma(-0.5 * x * y, y, 1.5);
}

// Row Gram-Schmidt of all three rows in FP64: the nearest unitary matrix
// in the row-by-row sense, with det U's phase kept (no SU(3) assumption).
template <bool FAST_NORM = false> __device__ __forceinline__ void orthonormalize_rows(M3<double> &m)
{
#pragma unroll
    for (int r = 0; r < 3; r++) {
#pragma unroll
        for (int q = 0; q < r; q++) {
            // row r -= <row q, row r> row q  (<a, b> = sum conj(a) b)
            double pr = 0.0, pi = 0.0;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                pr += m.e[q][j].r * m.e[r][j].r + m.e[q][j].i * m.e[r][j].i;
                pi += m.e[q][j].r * m.e[r][j].i - m.e[q][j].i * m.e[r][j].r;
            }
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const double ar = m.e[q][j].r, ai = m.e[q][j].i;
                m.e[r][j].r -= pr * ar - pi * ai;
                m.e[r][j].i -= pr * ai + pi * ar;
            }
        }
        double n = 0.0;
#pragma unroll
        for (int j = 0; j < 3; j++) n += m.e[r][j].r * m.e[r][j].r + m.e[r][j].i * m.e[r][j].i;
        n = FAST_NORM ? rsqrt_refined(n) : 1.0 / sqrt(n);
#pragma unroll
        for (int j = 0; j < 3; j++) {
            m.e[r][j].r *= n;
            m.e[r][j].i *= n;
        }
    }
}

// Row Gram-Schmidt in FP64; row 2 = conj(row0 x row1) so det = +1.
// FAST_NORM: the norms through rsqrt_refined (links stored in FP32).
template <int ROWS = 3, bool FAST_NORM = false> __device__ __forceinline__ void reunitarize(M3<double> &m)
{
    double n0 = 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++) n0 += m.e[0][j].r * m.e[0][j].r + m.e[0][j].i * m.e[0][j].i;
    n0 = FAST_NORM ? rsqrt_refined(n0) : 1.0 / sqrt(n0);
#pragma unroll
    for (int j = 0; j < 3; j++) {
        m.e[0][j].r *= n0;
        m.e[0][j].i *= n0;
    }
    double pr = 0.0, pi = 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        pr += m.e[0][j].r * m.e[1][j].r + m.e[0][j].i * m.e[1][j].i;
        pi += m.e[0][j].r * m.e[1][j].i - m.e[0][j].i * m.e[1][j].r;
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {
        double ar = m.e[0][j].r, ai = m.e[0][j].i;
        m.e[1][j].r -= pr * ar - pi * ai;
        m.e[1][j].i -= pr * ai + pi * ar;
    }
    double n1 = 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++) n1 += m.e[1][j].r * m.e[1][j].r + m.e[1][j].i * m.e[1][j].i;
    n1 = FAST_NORM ? rsqrt_refined(n1) : 1.0 / sqrt(n1);
#pragma unroll
    for (int j = 0; j < 3; j++) {
        m.e[1][j].r *= n1;
        m.e[1][j].i *= n1;
    }
    if constexpr (ROWS == 3) {  // ROWS == 2: the caller stores two rows only
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int k = (j + 1) % 3, l = (j + 2) % 3;
            cpx<double> x = cmul(m.e[0][k], m.e[1][l]);
            cpx<double> y = cmul(m.e[0][l], m.e[1][k]);
            m.e[2][j] = {x.r - y.r, -(x.i - y.i)};
        }
    }
}

// Reunitarisation for three stored rows that keeps the sign of det U: rows 0
// and 1 by Gram-Schmidt, row 2 = s conj(row0 x row1) with s = sign(Re det U)
// of the input.  det U = sum_j U_2j C_j with the cofactors
// C_j = U_0k U_1l - U_0l U_1k ((j, k, l) cyclic), and conj(C) is the third
// row of determinant +1, so s = -1 rebuilds a det = -1 link as det = -1.
// GLAC's generateRandom hot start makes such links and its SU(3) updates keep
// them (GLAC/math/matrices/su3matrixgenerator.h:156-161, GLAC/system.cpp:330);
// for det = +1 links this is reunitarize<3>.
template <bool FAST_NORM = false> __device__ __forceinline__ void reunitarize_keep_det_sign(M3<double> &m)
{
    reunitarize<2, FAST_NORM>(m);  // rows 0 and 1; row 2 still the input's
    cpx<double> cof[3];
    double det_re = 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int k = (j + 1) % 3, l = (j + 2) % 3;
        const cpx<double> x = cmul(m.e[0][k], m.e[1][l]);
        const cpx<double> y = cmul(m.e[0][l], m.e[1][k]);
        cof[j] = {x.r - y.r, x.i - y.i};
        det_re += m.e[2][j].r * cof[j].r - m.e[2][j].i * cof[j].i;
    }
    const double s = det_re < 0.0 ? -1.0 : 1.0;
#pragma unroll
    for (int j = 0; j < 3; j++) m.e[2][j] = {s * cof[j].r, -(s * cof[j].i)};
}

} // namespace glacx
