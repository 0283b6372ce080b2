/*
 * glac_oracle.c -- FP64 CPU restatement of GLAC's gauge-generation and
 * gradient-flow path (declarations and layout in glac_oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker / timed CPU baseline.  The
 * product never links or calls it.
 *
 * Each function names the reference lines it restates and the checks that
 * pin it (all under tests/test_oracle_*.py):
 *
 *  or_su3_mul / or_su3_dag / or_su3_retrace_mul
 *      GLAC/math/matrices/su3.h:184-221, 307-337; GLAC/math/functions.h:45-53.
 *      Pinned: GLAC's integer fixtures U1*U2 = UMul, U1^dagger = UCT,
 *      Re tr(U1 UTrace) = -21 (GLAC/tests/unit_tests/testcore.cpp:31-166).
 *  or_generate_su2 / or_rst_mul / or_rst_mul_inv / or_generate_rst
 *      GLAC/math/matrices/su3matrixgenerator.h:170-321.
 *      Pinned: U_RST known answer (testcore.cpp:111-135) and its inverse
 *      (randommatrixtests.cpp:10-46); SU(2) = x0 + i x.sigma closed form;
 *      RST output in SU(3); sign(r0) and inverse-coin statistics.
 *  or_generate_random  su3matrixgenerator.h:98-164.
 *      Pinned: unitarity, det = -1 (closed form of GLAC's third column),
 *      first column = normalised first input column.
 *  or_mt64_* / or_philox4x32_10 / or_xoshiro128ss_next
 *      C++11 [rand.predef] mt19937_64 (10000th output 9981545732273789042)
 *      and libstdc++ uniform_real_distribution (compared with g++'s <random>);
 *      Philox4x32-10 of Salmon et al., SC'11 (Random123 known-answer
 *      vectors); xoshiro128** 1.1 of Blackman & Vigna (first ten outputs
 *      from the state (1, 2, 3, 4) of their reference code).
 *  or_staple / or_delta_action   GLAC/actions/wilsongaugeaction.cpp:9, 24-53.
 *      Pinned: unit field with U' = 2*1, beta = 6 gives dS = -36 at every
 *      link (GLAC/tests/unit_tests/actiontests.cpp:67-114); staple of a pure
 *      gauge field; gauge covariance; dS against the brute-force change of
 *      the total Wilson action.
 *  or_action_derivative   GLAC/actions/wilsongaugeaction.cpp:55-86.
 *      Pinned: equals GLAC's explicit Gell-Mann projection
 *      (GLAC/actions/wilsonexplicitder.cpp:88-107, restated in the test);
 *      pure gauge gives Z = 0; finite-difference derivative of the Wilson
 *      action along exp(s T^a) U.
 *  or_su3_exp   GLAC/math/exponentiation/su3exp.cpp:16-79 (Morningstar &
 *      Peardon, PRD 69 054501).  Pinned: scipy.linalg.expm on random su(3),
 *      diagonal closed form, det = 1, unitarity, c0 < 0 branch, |w| = 0.05
 *      branch continuity.  Divergence from GLAC: GLAC returns NaN for Q = 0
 *      (0/0 in acos); here Q -> 0 returns the Taylor series.
 *  or_flow_step   GLAC/flow/flow.cpp:43-73.
 *      Pinned: third-order global convergence against a fine-step solution
 *      (a wrong RK3 coefficient breaks the order), pure gauge stays pure
 *      gauge, E(t) non-increasing, gauge invariance along the flow.
 *  or_measure   GLAC/observables/mastersamplertopcxyz.cpp:236-384.
 *      Pinned: unit field (plaq = 1, E = Q = 0); constant abelian fields
 *      with closed-form plaquette and energy and integer charge Q; gauge
 *      invariance; sum_t Q(t) = Q; parity flips Q.
 *  or_measure_density   GLAC/observables/latticeactionchargedensity.cpp:
 *      169-306.  Pinned: sums equal or_measure's Q and V E_raw; every site's
 *      values are gauge invariant; a lattice translation of the field
 *      translates the densities; time reflection flips the charge density.
 *  or_measure_super   GLAC/observables/supersampler.cpp:298-496 (adds the
 *      Weinberg operator W and W(t) to or_measure's observables).  Pinned:
 *      plaq, E, Q, Q(t) equal or_measure's; W = 0 on commuting (abelian)
 *      fields with flux in three planes sharing a direction, where a wrong
 *      orientation of one G_{nu lambda} term would leave a non-zero
 *      commutator trace; gauge invariance; time reflection flips W;
 *      sum_t W(t) = W; W scales as s^3 for U = exp(i s H), s -> 0 (three
 *      field-strength factors).  The overall scale 2 * 8 * 8 / (3 * 16^3) is GLAC's
 *      constant (supersampler.cpp:19) and parity unpinned beyond that
 *      reading: GLAC ships no Weinberg reference values.
 *  or_glac_hot_start / or_glac_sweep   GLAC/system.cpp:200-214, 321-361.
 *      Pinned statistically: strong-coupling plaquette beta/18, acceptance
 *      near GLAC's recorded 0.108 at 4^3x8, beta 6, eps 0.24, 10 hits
 *      (examples/test_config/configExampleLoad.out:72).  GLAC's exact chain is
 *      parity unpinned (GLAC needs MPI and is not built here).
 *  or_reunitarize / or_reunitarize_keep_det_sign / or_hot_start_haar /
 *  or_cb_sweep   no GLAC counterpart.
 *      Pinned: unitarity and det = 1; Haar moments <|tr U|^2> = 1;
 *      single-link histogram against exp(beta/3 Re tr UA); same plaquette
 *      distribution as or_glac_sweep; GLAC's recorded acceptance and
 *      plaquette (configExampleLoad.out:58-73); det = -1 kept by
 *      or_reunitarize_keep_det_sign.
 *  or_hot_start   GLAC/system.cpp:200-214 on the link streams.  Pinned: the
 *      GLAC_RANDOM and RST kinds have the moments of GLAC's own generators
 *      on its mt19937_64 stream (or_glac_hot_start), det = -1 / +1,
 *      unitarity; the HAAR kind is or_hot_start_haar.
 */
#include "glac_oracle.h"

#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdlib.h>
#include <string.h>

#define RE(m, i, j) ((m)[6 * (i) + 2 * (j)])
#define IM(m, i, j) ((m)[6 * (i) + 2 * (j) + 1])

/* Threads the parallel site loops use (OpenMP; 1 in a build without it). */
int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ======================================================================== */
/* 3x3 complex algebra                                                       */
/* ======================================================================== */

void or_su3_mul(const double *a, const double *b, double *c)
{
    double t[18];
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) {
            double re = 0.0, im = 0.0;
            for (int k = 0; k < 3; k++) {
                re += RE(a, i, k) * RE(b, k, j) - IM(a, i, k) * IM(b, k, j);
                im += RE(a, i, k) * IM(b, k, j) + IM(a, i, k) * RE(b, k, j);
            }
            t[6 * i + 2 * j] = re;
            t[6 * i + 2 * j + 1] = im;
        }
    }
    memcpy(c, t, sizeof t);
}

void or_su3_dag(const double *a, double *c)
{
    double t[18];
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) {
            t[6 * i + 2 * j] = RE(a, j, i);
            t[6 * i + 2 * j + 1] = -IM(a, j, i);
        }
    }
    memcpy(c, t, sizeof t);
}

double or_su3_retrace_mul(const double *a, const double *b)
{
    double s = 0.0;
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++)
            s += RE(a, i, k) * RE(b, k, i) - IM(a, i, k) * IM(b, k, i);
    return s;
}

void or_su3_det(const double *a, double *re_im)
{
    /* cofactor expansion along row 0 */
    double dr = 0.0, di = 0.0;
    for (int j = 0; j < 3; j++) {
        int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        /* minor = a1j1 a2j2 - a1j2 a2j1 (cyclic order gives the sign) */
        double mr = RE(a, 1, j1) * RE(a, 2, j2) - IM(a, 1, j1) * IM(a, 2, j2)
                  - (RE(a, 1, j2) * RE(a, 2, j1) - IM(a, 1, j2) * IM(a, 2, j1));
        double mi = RE(a, 1, j1) * IM(a, 2, j2) + IM(a, 1, j1) * RE(a, 2, j2)
                  - (RE(a, 1, j2) * IM(a, 2, j1) + IM(a, 1, j2) * RE(a, 2, j1));
        dr += RE(a, 0, j) * mr - IM(a, 0, j) * mi;
        di += RE(a, 0, j) * mi + IM(a, 0, j) * mr;
    }
    re_im[0] = dr;
    re_im[1] = di;
}

static void su3_zero(double *m) { memset(m, 0, 18 * sizeof(double)); }
static void su3_unit(double *m)
{
    su3_zero(m);
    m[0] = m[8] = m[16] = 1.0;
}
static void su3_add(double *a, const double *b, double s)
{
    for (int k = 0; k < 18; k++) a[k] += s * b[k];
}

/* ======================================================================== */
/* Random numbers                                                            */
/* ======================================================================== */

/* mt19937_64 with the parameters fixed by C++11 [rand.predef]. */
#define MT_N 312
#define MT_M 156
void or_mt64_seed(or_mt64 *g, uint64_t seed)
{
    g->mt[0] = seed;
    for (int i = 1; i < MT_N; i++)
        g->mt[i] = 6364136223846793005ULL * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
    g->mti = MT_N;
}

uint64_t or_mt64_next(or_mt64 *g)
{
    const uint64_t upper = 0xFFFFFFFF80000000ULL, lower = 0x7FFFFFFFULL;
    const uint64_t a = 0xB5026F5AA96619E9ULL;
    if (g->mti >= MT_N) {
        for (int i = 0; i < MT_N; i++) {
            uint64_t x = (g->mt[i] & upper) | (g->mt[(i + 1) % MT_N] & lower);
            uint64_t xa = x >> 1;
            if (x & 1ULL) xa ^= a;
            g->mt[i] = g->mt[(i + MT_M) % MT_N] ^ xa;
        }
        g->mti = 0;
    }
    uint64_t y = g->mt[g->mti++];
    y ^= (y >> 29) & 0x5555555555555555ULL;
    y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
    y ^= (y << 37) & 0xFFF7EEE000000000ULL;
    y ^= y >> 43;
    return y;
}

/* libstdc++ uniform_real_distribution<double>(a, b) over a 64-bit engine:
 * generate_canonical takes one draw, scales it by 2^-64 and clamps below 1. */
double or_mt64_uniform(or_mt64 *g, double a, double b)
{
    double c = (double)or_mt64_next(g) * 0x1p-64;
    if (c >= 1.0) c = nextafter(1.0, 0.0);
    return c * (b - a) + a;
}

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline uint32_t rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

/* xoshiro128** 1.1 (Blackman & Vigna). */
uint32_t or_xoshiro128ss_next(uint32_t s[4])
{
    uint32_t result = rotl32(s[1] * 5u, 7) * 9u;
    uint32_t t = s[1] << 9;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl32(s[3], 11);
    return result;
}

/* Counter-based per-link stream shared with the device kernels: one
 * Philox4x32-10 block keyed by the 64-bit seed with counter
 * (site, mu | tag << 8, lo32(sweep), hi32(sweep)) seeds a xoshiro128** state. */
enum { STREAM_METROPOLIS = 1, STREAM_HOT_START = 2 };
static void link_stream(uint32_t st[4], uint64_t seed, long site, int mu, int tag, uint64_t sweep)
{
    uint32_t ctr[4] = {(uint32_t)site, (uint32_t)mu | ((uint32_t)tag << 8), (uint32_t)sweep,
                       (uint32_t)(sweep >> 32)};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    or_philox4x32_10(ctr, key, st);
    if ((st[0] | st[1] | st[2] | st[3]) == 0u) st[0] = 1u;
}
/* uniform in (0, 1) with 32 bits */
static inline double stream_open(uint32_t s[4])
{
    return ((double)or_xoshiro128ss_next(s) + 0.5) * 0x1p-32;
}
/* two uniforms on the 2^16 midpoints of [-1/2, 1/2) from one output: lo from
 * its low, hi from its high 16 bits */
static inline void stream_centered_pair(uint32_t s[4], double *lo, double *hi)
{
    const uint32_t b = or_xoshiro128ss_next(s);
    *lo = (double)(b & 0xFFFFu) * 0x1p-16 - (0.5 - 0x1p-17);
    *hi = (double)(b >> 16) * 0x1p-16 - (0.5 - 0x1p-17);
}
/* GLAC's 13 RST uniforms (r, s, t, coin) as the 16-bit halves of seven outputs */
static inline void stream_rst_uniforms(uint32_t s[4], double u[13])
{
    double q[14];
    for (int k = 0; k < 7; k++) stream_centered_pair(s, &q[2 * k], &q[2 * k + 1]);
    memcpy(u, q, 13 * sizeof(double));
}
/* uniform in (-1, 1) with 32 bits */
static inline double stream_symmetric(uint32_t s[4])
{
    return ((double)or_xoshiro128ss_next(s) + 0.5) * 0x1p-31 - 1.0;
}

/* ======================================================================== */
/* GLAC random SU(3) generators                                              */
/* ======================================================================== */

/* SU(2) near unity: x0 = sign(r0) sqrt(1 - eps^2), x_k = eps r_k / |r|,
 * renormalised, U = x0 1 + i x.sigma, 2x2 complex row-major interleaved. */
void or_generate_su2(const double r[4], double eps, double out[8])
{
    double x[4];
    double x0 = sqrt(1.0 - eps * eps);
    x[0] = (r[0] < 0.0) ? -x0 : x0;
    double rn = sqrt(r[1] * r[1] + r[2] * r[2] + r[3] * r[3]) / eps;
    for (int k = 1; k < 4; k++) x[k] = r[k] / rn;
    double nrm = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + x[3] * x[3]);
    for (int k = 0; k < 4; k++) x[k] /= nrm;
    /* sigma_1 = [[0,1],[1,0]], sigma_2 = [[0,-i],[i,0]], sigma_3 = diag(1,-1) */
    out[0] = x[0];  out[1] = x[3];   /* (0,0) = x0 + i x3 */
    out[2] = x[2];  out[3] = x[1];   /* (0,1) = x2 + i x1 */
    out[4] = -x[2]; out[5] = x[1];   /* (1,0) = -x2 + i x1 */
    out[6] = x[0];  out[7] = -x[3];  /* (1,1) = x0 - i x3 */
}

/* Embed a 2x2 complex matrix into rows/cols (p, q) of a 3x3 unit matrix. */
static void embed_su2(const double *s, int p, int q, double *M)
{
    su3_unit(M);
    int idx[2] = {p, q};
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            RE(M, idx[a], idx[b]) = s[4 * a + 2 * b];
            IM(M, idx[a], idx[b]) = s[4 * a + 2 * b + 1];
        }
}

/* X = R S T with r in block (0,1), s in block (0,2), t in block (1,2). */
void or_rst_mul(const double *r, const double *s, const double *t, double *X)
{
    double R[18], S[18], T[18];
    embed_su2(r, 0, 1, R);
    embed_su2(s, 0, 2, S);
    embed_su2(t, 1, 2, T);
    or_su3_mul(R, S, X);
    or_su3_mul(X, T, X);
}

void or_rst_mul_inv(const double *r, const double *s, const double *t, double *X)
{
    or_rst_mul(r, s, t, X);
    or_su3_dag(X, X);
}

/* u[0..3] -> r, u[4..7] -> s, u[8..11] -> t (each U(-1/2, 1/2)), u[12] < 0
 * selects the inverse product, in GLAC's draw order. */
void or_generate_rst(const double u[13], double eps, double X[18])
{
    double r[8], s[8], t[8];
    or_generate_su2(u + 0, eps, r);
    or_generate_su2(u + 4, eps, s);
    or_generate_su2(u + 8, eps, t);
    if (u[12] < 0.0)
        or_rst_mul_inv(r, s, t, X);
    else
        or_rst_mul(r, s, t, X);
}

/* GLAC's "fully random" matrix: entries from the uniforms (row-major, re then
 * im), Gram-Schmidt on columns 0 and 1, column 2 = -conj(c0 x c1), which makes
 * det = -1. */
void or_generate_random(const double u[18], double H[18])
{
    memcpy(H, u, 18 * sizeof(double));
    double n0 = 0.0;
    for (int i = 0; i < 3; i++) n0 += RE(H, i, 0) * RE(H, i, 0) + IM(H, i, 0) * IM(H, i, 0);
    n0 = sqrt(n0);
    for (int i = 0; i < 3; i++) { RE(H, i, 0) /= n0; IM(H, i, 0) /= n0; }
    double pr = 0.0, pi = 0.0; /* c0^dagger c1 */
    for (int i = 0; i < 3; i++) {
        pr += RE(H, i, 0) * RE(H, i, 1) + IM(H, i, 0) * IM(H, i, 1);
        pi += RE(H, i, 0) * IM(H, i, 1) - IM(H, i, 0) * RE(H, i, 1);
    }
    for (int i = 0; i < 3; i++) {
        double cr = RE(H, i, 0), ci = IM(H, i, 0);
        RE(H, i, 1) -= cr * pr - ci * pi;
        IM(H, i, 1) -= cr * pi + ci * pr;
    }
    double n1 = 0.0;
    for (int i = 0; i < 3; i++) n1 += RE(H, i, 1) * RE(H, i, 1) + IM(H, i, 1) * IM(H, i, 1);
    n1 = sqrt(n1);
    for (int i = 0; i < 3; i++) { RE(H, i, 1) /= n1; IM(H, i, 1) /= n1; }
    for (int i = 0; i < 3; i++) {
        int j = (i + 1) % 3, k = (i + 2) % 3;
        /* (c0 x c1)_i = c0_j c1_k - c0_k c1_j */
        double xr = RE(H, j, 0) * RE(H, k, 1) - IM(H, j, 0) * IM(H, k, 1)
                  - (RE(H, k, 0) * RE(H, j, 1) - IM(H, k, 0) * IM(H, j, 1));
        double xi = RE(H, j, 0) * IM(H, k, 1) + IM(H, j, 0) * RE(H, k, 1)
                  - (RE(H, k, 0) * IM(H, j, 1) + IM(H, k, 0) * RE(H, j, 1));
        RE(H, i, 2) = -xr;
        IM(H, i, 2) = xi;
    }
}

/* ======================================================================== */
/* Lattice geometry                                                          */
/* ======================================================================== */

static inline long volume(const int n[4]) { return (long)n[0] * n[1] * n[2] * n[3]; }

static inline long lex(const int n[4], const int c[4])
{
    return c[0] + (long)n[0] * (c[1] + (long)n[1] * (c[2] + (long)n[2] * c[3]));
}

/* site index of c + s1 e_{d1} + s2 e_{d2}, periodic */
static inline long shifted(const int n[4], const int c[4], int d1, int s1, int d2, int s2)
{
    int m[4] = {c[0], c[1], c[2], c[3]};
    m[d1] = (m[d1] + s1 + n[d1]) % n[d1];
    if (s2 != 0) m[d2] = (m[d2] + s2 + n[d2]) % n[d2];
    return lex(n, m);
}

static inline const double *link(const double *U, long V, int mu, long site)
{
    return U + ((long)mu * V + site) * 18;
}

static void decode(const int n[4], long site, int c[4])
{
    c[0] = (int)(site % n[0]); site /= n[0];
    c[1] = (int)(site % n[1]); site /= n[1];
    c[2] = (int)(site % n[2]); site /= n[2];
    c[3] = (int)site;
}

/* ======================================================================== */
/* Wilson action                                                             */
/* ======================================================================== */

static void staple_at(const double *U, const int n[4], const int c[4], int mu, double *A)
{
    long V = volume(n);
    long x = lex(n, c);
    double t1[18], t2[18], d[18];
    su3_zero(A);
    for (int nu = 0; nu < 4; nu++) {
        if (nu == mu) continue;
        /* upper: U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag */
        long xpm = shifted(n, c, mu, +1, nu, 0);
        long xpn = shifted(n, c, nu, +1, mu, 0);
        or_su3_dag(link(U, V, mu, xpn), d);
        or_su3_mul(link(U, V, nu, xpm), d, t1);
        or_su3_dag(link(U, V, nu, x), d);
        or_su3_mul(t1, d, t1);
        /* lower: U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu) */
        long xpmmn = shifted(n, c, mu, +1, nu, -1);
        long xmn = shifted(n, c, nu, -1, mu, 0);
        or_su3_dag(link(U, V, nu, xpmmn), t2);
        or_su3_dag(link(U, V, mu, xmn), d);
        or_su3_mul(t2, d, t2);
        or_su3_mul(t2, link(U, V, nu, xmn), t2);
        su3_add(A, t1, 1.0);
        su3_add(A, t2, 1.0);
    }
}

void or_staple(const double *U, const int n[4], int x, int y, int z, int t, int mu, double A[18])
{
    int c[4] = {x, y, z, t};
    staple_at(U, n, c, mu, A);
}

double or_delta_action(const double *Uold, const double *Unew, const double *A, double beta)
{
    double D[18];
    for (int k = 0; k < 18; k++) D[k] = Unew[k] - Uold[k];
    return -(beta / 3.0) * or_su3_retrace_mul(D, A);
}

/* Z = 1/2 [ (O^dag - O) - 1/3 Im tr(O^dag - O) i 1 ],  O = U_mu A_mu. */
static void force_from_omega(const double *O, double *Z)
{
    double Od[18];
    or_su3_dag(O, Od);
    for (int k = 0; k < 18; k++) Z[k] = Od[k] - O[k];
    double itr = (Z[1] + Z[9] + Z[17]) / 3.0;
    Z[1] -= itr;
    Z[9] -= itr;
    Z[17] -= itr;
    for (int k = 0; k < 18; k++) Z[k] *= 0.5;
}

void or_action_derivative(const double *U, const int n[4], int mu, double *Z)
{
    long V = volume(n);
#pragma omp parallel for schedule(static)
    for (long s = 0; s < V; s++) {
        int c[4];
        double A[18], O[18];
        decode(n, s, c);
        staple_at(U, n, c, mu, A);
        or_su3_mul(link(U, V, mu, s), A, O);
        force_from_omega(O, Z + s * 18);
    }
}

/* ======================================================================== */
/* SU(3) exponential                                                         */
/* ======================================================================== */

void or_su3_exp_taylor(const double *Q, int order, double *out)
{
    double term[18], acc[18];
    su3_unit(term);
    su3_unit(acc);
    for (int k = 1; k <= order; k++) {
        or_su3_mul(term, Q, term);
        for (int j = 0; j < 18; j++) term[j] /= (double)k;
        su3_add(acc, term, 1.0);
    }
    memcpy(out, acc, sizeof acc);
}

/* exp(Q) for traceless antihermitian Q via Cayley-Hamilton:
 * H = -iQ, exp(Q) = exp(iH) = f0 1 + f1 H + f2 H^2 with the f_j of
 * Morningstar & Peardon (PRD 69 054501, eqs. 29-34). */
void or_su3_exp(const double *Q, double *out)
{
    double H[18], H2[18], H3[18];
    for (int k = 0; k < 9; k++) {
        H[2 * k] = Q[2 * k + 1];
        H[2 * k + 1] = -Q[2 * k];
    }
    or_su3_mul(H, H, H2);
    or_su3_mul(H, H2, H3);
    double c0 = (H3[0] + H3[8] + H3[16]) / 3.0;
    double c1 = 0.5 * (H2[0] + H2[8] + H2[16]);
    if (!(c1 > 1e-28)) { /* |Q| < 1e-14: the series is exact to rounding */
        or_su3_exp_taylor(Q, 3, out);
        return;
    }
    double c0max = 2.0 * (c1 / 3.0) * sqrt(c1 / 3.0);
    double ratio = fabs(c0) / c0max;
    if (ratio > 1.0) ratio = 1.0;
    double theta = acos(ratio);
    double u = sqrt(c1 / 3.0) * cos(theta / 3.0);
    double w = sqrt(c1) * sin(theta / 3.0);
    double uu = u * u, ww = w * w;
    double xi0 = (fabs(w) > 0.05) ? sin(w) / w
                                  : 1.0 - ww / 6.0 * (1.0 - ww / 20.0 * (1.0 - ww / 42.0));
    double cu = cos(u), su = sin(u), cw = cos(w), c2u = cos(2.0 * u), s2u = sin(2.0 * u);
    /* e^{2iu} = c2u + i s2u ; e^{-iu} = cu - i su */
    double h[3][2];
    {
        /* h0 = (u^2 - w^2) e^{2iu} + e^{-iu} [8u^2 cos w + 2iu(3u^2 + w^2) xi0] */
        double br = 8.0 * uu * cw, bi = 2.0 * u * (3.0 * uu + ww) * xi0;
        h[0][0] = (uu - ww) * c2u + (cu * br + su * bi);
        h[0][1] = (uu - ww) * s2u + (cu * bi - su * br);
    }
    {
        /* h1 = 2u e^{2iu} - e^{-iu} [2u cos w - i(3u^2 - w^2) xi0] */
        double br = 2.0 * u * cw, bi = -(3.0 * uu - ww) * xi0;
        h[1][0] = 2.0 * u * c2u - (cu * br + su * bi);
        h[1][1] = 2.0 * u * s2u - (cu * bi - su * br);
    }
    {
        /* h2 = e^{2iu} - e^{-iu} [cos w + 3iu xi0] */
        double br = cw, bi = 3.0 * u * xi0;
        h[2][0] = c2u - (cu * br + su * bi);
        h[2][1] = s2u - (cu * bi - su * br);
    }
    double den = 9.0 * uu - ww;
    double f[3][2];
    for (int j = 0; j < 3; j++) {
        f[j][0] = h[j][0] / den;
        f[j][1] = h[j][1] / den;
    }
    if (c0 < 0.0) { /* f_j(-c0) = (-1)^j conj(f_j(c0)) */
        f[0][1] = -f[0][1];
        f[1][0] = -f[1][0];
        f[2][1] = -f[2][1];
    }
    for (int k = 0; k < 9; k++) {
        double hr = H[2 * k], hi = H[2 * k + 1], qr = H2[2 * k], qi = H2[2 * k + 1];
        out[2 * k] = f[1][0] * hr - f[1][1] * hi + f[2][0] * qr - f[2][1] * qi;
        out[2 * k + 1] = f[1][0] * hi + f[1][1] * hr + f[2][0] * qi + f[2][1] * qr;
    }
    for (int d = 0; d < 3; d++) {
        out[8 * d] += f[0][0];
        out[8 * d + 1] += f[0][1];
    }
}

/* ======================================================================== */
/* RK3 Wilson flow (Luscher, JHEP 08 (2010) 071)                             */
/* ======================================================================== */

static void exp_times_field(double *U, long V, const double *Y, double scale)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < 4 * V; i++) {
        double Q[18], X[18];
        for (int k = 0; k < 18; k++) Q[k] = scale * Y[i * 18 + k];
        or_su3_exp(Q, X);
        or_su3_mul(X, U + i * 18, U + i * 18);
    }
}

void or_flow_step(double *U, const int n[4], double eps)
{
    long V = volume(n);
    double *Z0 = (double *)malloc((size_t)4 * V * 18 * sizeof(double));
    double *Z = (double *)malloc((size_t)4 * V * 18 * sizeof(double));
    /* W1 = exp(eps/4 Z0) W0 */
    for (int mu = 0; mu < 4; mu++) or_action_derivative(U, n, mu, Z0 + (long)mu * V * 18);
    exp_times_field(U, V, Z0, 0.25 * eps);
    /* Y = 8/9 eps Z(W1) - 17/36 eps Z0 ;  W2 = exp(Y) W1 */
    for (int mu = 0; mu < 4; mu++) or_action_derivative(U, n, mu, Z + (long)mu * V * 18);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < 4 * V * 18; i++) Z0[i] = (8.0 / 9.0) * eps * Z[i] - (17.0 / 36.0) * eps * Z0[i];
    exp_times_field(U, V, Z0, 1.0);
    /* Y' = 3/4 eps Z(W2) - Y ;  W3 = exp(Y') W2 */
    for (int mu = 0; mu < 4; mu++) or_action_derivative(U, n, mu, Z + (long)mu * V * 18);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < 4 * V * 18; i++) Z0[i] = 0.75 * eps * Z[i] - Z0[i];
    exp_times_field(U, V, Z0, 1.0);
    free(Z0);
    free(Z);
}

/* ======================================================================== */
/* Clover observables                                                        */
/* ======================================================================== */

/* Four-leaf clover in the (mu, nu) plane with GLAC's leaf orientations:
 * C = L1 - L2 + L3 - L4;  also returns Re tr L1 (the plaquette). */
static double clover(const double *U, const int n[4], const int c[4], int mu, int nu, double *C)
{
    long V = volume(n);
    long x = lex(n, c);
    long xpm = shifted(n, c, mu, +1, nu, 0), xpn = shifted(n, c, nu, +1, mu, 0);
    long xmm = shifted(n, c, mu, -1, nu, 0), xmn = shifted(n, c, nu, -1, mu, 0);
    long xpmmn = shifted(n, c, mu, +1, nu, -1), xmmmn = shifted(n, c, mu, -1, nu, -1);
    long xmmpn = shifted(n, c, mu, -1, nu, +1);
    double L[18], d[18];
    /* L1 = U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag */
    or_su3_mul(link(U, V, mu, x), link(U, V, nu, xpm), L);
    or_su3_dag(link(U, V, mu, xpn), d); or_su3_mul(L, d, L);
    or_su3_dag(link(U, V, nu, x), d);   or_su3_mul(L, d, L);
    double plaq = L[0] + L[8] + L[16];
    memcpy(C, L, sizeof L);
    /* L2 = U_mu(x) U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu) */
    or_su3_dag(link(U, V, nu, xpmmn), d); or_su3_mul(link(U, V, mu, x), d, L);
    or_su3_dag(link(U, V, mu, xmn), d);   or_su3_mul(L, d, L);
    or_su3_mul(L, link(U, V, nu, xmn), L);
    su3_add(C, L, -1.0);
    /* L3 = U_mu(x-mu)^dag U_nu(x-mu-nu)^dag U_mu(x-mu-nu) U_nu(x-nu) */
    or_su3_dag(link(U, V, mu, xmm), L);
    or_su3_dag(link(U, V, nu, xmmmn), d); or_su3_mul(L, d, L);
    or_su3_mul(L, link(U, V, mu, xmmmn), L);
    or_su3_mul(L, link(U, V, nu, xmn), L);
    su3_add(C, L, 1.0);
    /* L4 = U_mu(x-mu)^dag U_nu(x-mu) U_mu(x-mu+nu) U_nu(x)^dag */
    or_su3_dag(link(U, V, mu, xmm), L);
    or_su3_mul(L, link(U, V, nu, xmm), L);
    or_su3_mul(L, link(U, V, mu, xmmpn), L);
    or_su3_dag(link(U, V, nu, x), d); or_su3_mul(L, d, L);
    su3_add(C, L, -1.0);
    return plaq;
}

/* C <- traceless(C - C^dag) */
static void traceless_antiherm(double *C)
{
    double Cd[18];
    or_su3_dag(C, Cd);
    for (int k = 0; k < 18; k++) C[k] -= Cd[k];
    double itr = (C[1] + C[9] + C[17]) / 3.0;
    C[1] -= itr;
    C[9] -= itr;
    C[17] -= itr;
}

/* out = {plaquette, E_raw, Q};  qt[t] = Q summed over x, y, z.
 * plaquette = sum_x sum_{mu<nu} Re tr P / (18 V), E_raw = sum_x sum_{mu<nu}
 * Re tr(C C) / V (physical E = -E_raw / 64), Q = -sum_x sum_{3 pairs}
 * Re tr(C_{0 nu} C_{rho sigma}) / (256 pi^2) with (nu, rho, sigma) =
 * (1,2,3), (2,3,1), (3,1,2). */
void or_measure(const double *U, const int n[4], double out[3], double *qt)
{
    long V = volume(n);
    /* per-site terms in parallel (10 per site: 6 plaquettes, 3 energy pairs,
     * the charge), then summed serially in site order, so the result does
     * not depend on the thread count */
    double *terms = (double *)malloc((size_t)V * 10 * sizeof(double));
#pragma omp parallel for schedule(static)
    for (long s = 0; s < V; s++) {
        int c[4];
        decode(n, s, c);
        double *ts = terms + s * 10;
        double qs = 0.0;
        for (int nu = 1; nu < 4; nu++) {
            int rho = nu % 3 + 1, sigma = rho % 3 + 1;
            double C1[18], C2[18];
            ts[2 * (nu - 1)] = clover(U, n, c, 0, nu, C1);
            ts[2 * (nu - 1) + 1] = clover(U, n, c, rho, sigma, C2);
            traceless_antiherm(C1);
            traceless_antiherm(C2);
            qs -= or_su3_retrace_mul(C1, C2);
            ts[6 + nu - 1] = or_su3_retrace_mul(C1, C1) + or_su3_retrace_mul(C2, C2);
        }
        ts[9] = qs;
    }
    double plaq = 0.0, energy = 0.0;
    for (int t = 0; t < n[3]; t++) qt[t] = 0.0;
    for (long s = 0; s < V; s++) {
        const double *ts = terms + s * 10;
        for (int k = 0; k < 6; k++) plaq += ts[k];
        for (int k = 6; k < 9; k++) energy += ts[k];
        qt[(int)(s / ((long)n[0] * n[1] * n[2]))] += ts[9];
    }
    free(terms);
    const double qnorm = 1.0 / (256.0 * M_PI * M_PI);
    double q = 0.0;
    for (int t = 0; t < n[3]; t++) {
        qt[t] *= qnorm;
        q += qt[t];
    }
    out[0] = plaq / (18.0 * (double)V);
    out[1] = energy / (double)V;
    out[2] = q;
}

/* Per-site densities of GLAC's LatticeActionChargeDensity
 * (GLAC/observables/latticeactionchargedensity.cpp:169-306), lexicographic
 * site order: qd = -sum_nu Re tr(G_{0 nu} G_{rho sigma}) / (256 pi^2),
 * ed = sum_nu [Re tr(G_{0 nu}^2) + Re tr(G_{rho sigma}^2)]. */
void or_measure_density(const double *U, const int n[4], double *qd, double *ed)
{
    long V = volume(n);
    const double qnorm = 1.0 / (256.0 * M_PI * M_PI);
#pragma omp parallel for schedule(static)
    for (long s = 0; s < V; s++) {
        int c[4];
        decode(n, s, c);
        double qs = 0.0, es = 0.0;
        for (int nu = 1; nu < 4; nu++) {
            int rho = nu % 3 + 1, sigma = rho % 3 + 1;
            double C1[18], C2[18];
            clover(U, n, c, 0, nu, C1);
            clover(U, n, c, rho, sigma, C2);
            traceless_antiherm(C1);
            traceless_antiherm(C2);
            qs -= or_su3_retrace_mul(C1, C2);
            es += or_su3_retrace_mul(C1, C1);
            es += or_su3_retrace_mul(C2, C2);
        }
        qd[s] = qs * qnorm;
        ed[s] = es;
    }
}

/* GLAC's SuperSampler: the clover observables plus the Weinberg operator
 * W = -(2 * 8 * 8 / (3 * 16^3)) sum_x sum_{nu=1..3} sum_{lambda != 0, nu}
 *       Re tr(G_{0 lambda} G_{nu lambda} G_{rho sigma}),
 * (nu, rho, sigma) cyclic in (1, 2, 3), G = traceless(C - C^dag) of the
 * clover, G_{nu lambda} = G_{lambda nu}^dag for the orientations GLAC does
 * not store (supersampler.cpp:430-460, index map supersampler.h:65-70).
 * out = {plaq, E_raw, Q, W}; qt[Nt], wt[Nt] per time slice. */
void or_measure_super(const double *U, const int n[4], double out[4], double *qt, double *wt)
{
    long V = volume(n);
    double plaq = 0.0, energy = 0.0;
    for (int t = 0; t < n[3]; t++) qt[t] = wt[t] = 0.0;
    for (long s = 0; s < V; s++) {
        int c[4];
        decode(n, s, c);
        double G[4][4][18]; /* G[a][b] for a < b as computed, other order by dagger */
        double qs = 0.0, ws = 0.0;
        for (int nu = 1; nu < 4; nu++) {
            int rho = nu % 3 + 1, sigma = rho % 3 + 1;
            double C1[18], C2[18];
            plaq += clover(U, n, c, 0, nu, C1);
            plaq += clover(U, n, c, rho, sigma, C2);
            traceless_antiherm(C1);
            traceless_antiherm(C2);
            qs -= or_su3_retrace_mul(C1, C2);
            energy += or_su3_retrace_mul(C1, C1) + or_su3_retrace_mul(C2, C2);
            memcpy(G[0][nu], C1, sizeof C1);
            memcpy(G[rho][sigma], C2, sizeof C2);
            or_su3_dag(C1, G[nu][0]);
            or_su3_dag(C2, G[sigma][rho]);
        }
        for (int nu = 1; nu < 4; nu++) {
            int rho = nu % 3 + 1, sigma = rho % 3 + 1;
            for (int lam = 1; lam < 4; lam++) {
                if (lam == nu) continue;
                double T[18];
                or_su3_mul(G[0][lam], G[nu][lam], T);
                ws -= or_su3_retrace_mul(T, G[rho][sigma]);
            }
        }
        qt[c[3]] += qs;
        wt[c[3]] += ws;
    }
    const double qnorm = 1.0 / (256.0 * M_PI * M_PI), wnorm = 2.0 * 8.0 * 8.0 / (3.0 * 16.0 * 16.0 * 16.0);
    double q = 0.0, w = 0.0;
    for (int t = 0; t < n[3]; t++) {
        qt[t] *= qnorm;
        wt[t] *= wnorm;
        q += qt[t];
        w += wt[t];
    }
    out[0] = plaq / (18.0 * (double)V);
    out[1] = energy / (double)V;
    out[2] = q;
    out[3] = w;
}

/* ======================================================================== */
/* Metropolis, GLAC's sequential order and RNG streams                       */
/* ======================================================================== */

or_glac_rng *or_glac_rng_new(uint64_t metropolis_seed, uint64_t matrix_seed)
{
    or_glac_rng *g = (or_glac_rng *)malloc(sizeof(or_glac_rng));
    or_mt64_seed(&g->metro, metropolis_seed);
    or_mt64_seed(&g->matrix, matrix_seed);
    return g;
}

void or_glac_rng_free(or_glac_rng *g) { free(g); }

static void glac_rst(or_glac_rng *g, double eps, double *X)
{
    double u[13];
    for (int k = 0; k < 13; k++) u[k] = or_mt64_uniform(&g->matrix, -0.5, 0.5);
    or_generate_rst(u, eps, X);
}

void or_glac_hot_start(double *U, const int n[4], or_glac_rng *g, double eps, int rst)
{
    long V = volume(n);
    for (int mu = 0; mu < 4; mu++)
        for (long s = 0; s < V; s++) {
            double *m = U + ((long)mu * V + s) * 18;
            if (rst) {
                glac_rst(g, eps, m);
            } else {
                double u[18];
                for (int k = 0; k < 18; k++) u[k] = or_mt64_uniform(&g->matrix, -eps, eps);
                or_generate_random(u, m);
            }
        }
}

double or_glac_sweep(double *U, const int n[4], or_glac_rng *g, double beta, double eps, int nhit)
{
    long V = volume(n);
    unsigned long long accepted = 0;
    for (int x = 0; x < n[0]; x++)
        for (int y = 0; y < n[1]; y++)
            for (int z = 0; z < n[2]; z++)
                for (int t = 0; t < n[3]; t++)
                    for (int mu = 0; mu < 4; mu++) {
                        int c[4] = {x, y, z, t};
                        double A[18], X[18], Up[18];
                        double *Ul = U + ((long)mu * V + lex(n, c)) * 18;
                        staple_at(U, n, c, mu, A);
                        for (int h = 0; h < nhit; h++) {
                            glac_rst(g, eps, X);
                            or_su3_mul(X, Ul, Up);
                            double dS = or_delta_action(Ul, Up, A, beta);
                            if (exp(-dS) > or_mt64_uniform(&g->metro, 0.0, 1.0)) {
                                memcpy(Ul, Up, sizeof Up);
                                accepted++;
                            }
                        }
                    }
    return (double)accepted / (4.0 * nhit * (double)V);
}

/* ======================================================================== */
/* Metropolis, checkerboard order with the counter-based link streams        */
/* ======================================================================== */

/* Row Gram-Schmidt; row 2 = conj(row0 x row1), so det = +1. */
void or_reunitarize(double *m)
{
    double n0 = 0.0;
    for (int j = 0; j < 3; j++) n0 += RE(m, 0, j) * RE(m, 0, j) + IM(m, 0, j) * IM(m, 0, j);
    n0 = 1.0 / sqrt(n0);
    for (int j = 0; j < 3; j++) { RE(m, 0, j) *= n0; IM(m, 0, j) *= n0; }
    double pr = 0.0, pi = 0.0; /* sum_j conj(r0_j) r1_j */
    for (int j = 0; j < 3; j++) {
        pr += RE(m, 0, j) * RE(m, 1, j) + IM(m, 0, j) * IM(m, 1, j);
        pi += RE(m, 0, j) * IM(m, 1, j) - IM(m, 0, j) * RE(m, 1, j);
    }
    for (int j = 0; j < 3; j++) {
        double ar = RE(m, 0, j), ai = IM(m, 0, j);
        RE(m, 1, j) -= pr * ar - pi * ai;
        IM(m, 1, j) -= pr * ai + pi * ar;
    }
    double n1 = 0.0;
    for (int j = 0; j < 3; j++) n1 += RE(m, 1, j) * RE(m, 1, j) + IM(m, 1, j) * IM(m, 1, j);
    n1 = 1.0 / sqrt(n1);
    for (int j = 0; j < 3; j++) { RE(m, 1, j) *= n1; IM(m, 1, j) *= n1; }
    for (int j = 0; j < 3; j++) {
        int k = (j + 1) % 3, l = (j + 2) % 3;
        double xr = RE(m, 0, k) * RE(m, 1, l) - IM(m, 0, k) * IM(m, 1, l)
                  - (RE(m, 0, l) * RE(m, 1, k) - IM(m, 0, l) * IM(m, 1, k));
        double xi = RE(m, 0, k) * IM(m, 1, l) + IM(m, 0, k) * RE(m, 1, l)
                  - (RE(m, 0, l) * IM(m, 1, k) + IM(m, 0, l) * RE(m, 1, k));
        RE(m, 2, j) = xr;
        IM(m, 2, j) = -xi;
    }
}

/* Reunitarisation of a Metropolis-updated link that keeps the sign of det U:
 * rows 0 and 1 as or_reunitarize, row 2 = s conj(row0 x row1) with
 * s = sign(Re det U) of the input, det U = sum_j U_2j C_j (C_j the cofactors
 * of row 2, conj(C) the det = +1 third row).  GLAC's generateRandom start
 * gives det = -1 links (GLAC/math/matrices/su3matrixgenerator.h:156-161) and
 * its SU(3) updates keep them; for det = +1 this is or_reunitarize. */
void or_reunitarize_keep_det_sign(double *m)
{
    double row2[6];
    memcpy(row2, m + 12, sizeof row2);
    or_reunitarize(m);
    double det_re = 0.0;
    for (int j = 0; j < 3; j++) det_re += row2[2 * j] * RE(m, 2, j) + row2[2 * j + 1] * IM(m, 2, j);
    if (det_re < 0.0)
        for (int j = 0; j < 3; j++) { RE(m, 2, j) = -RE(m, 2, j); IM(m, 2, j) = -IM(m, 2, j); }
}

/* Hot starts from the link streams (seed, site, mu, tag 2, sweep 0):
 *  OR_HOT_HAAR: Haar-random SU(3), 9 complex Gaussians (Box-Muller on 18
 *      stream uniforms, entry k from uniforms 2k, 2k+1), then or_reunitarize;
 *  OR_HOT_GLAC_RANDOM: GLAC's generateRandom (su3matrixgenerator.h:98-164) on
 *      18 uniforms eps * stream_symmetric, entry k from uniforms 2k (re),
 *      2k+1 (im): det = -1;
 *  OR_HOT_RST: GLAC's generateRST (su3matrixgenerator.h:170-196) with step
 *      eps on the 13 uniforms of stream_rst_uniforms (GLAC/system.cpp:208). */
void or_hot_start(double *U, const int n[4], uint64_t seed, int kind, double eps)
{
    long V = volume(n);
    for (int mu = 0; mu < 4; mu++)
        for (long s = 0; s < V; s++) {
            uint32_t st[4];
            link_stream(st, seed, s, mu, STREAM_HOT_START, 0);
            double *m = U + ((long)mu * V + s) * 18;
            if (kind == OR_HOT_GLAC_RANDOM) {
                double u[18];
                for (int k = 0; k < 18; k++) u[k] = eps * stream_symmetric(st);
                or_generate_random(u, m);
            } else if (kind == OR_HOT_RST) {
                double u[13];
                stream_rst_uniforms(st, u);
                or_generate_rst(u, eps, m);
            } else {
                for (int k = 0; k < 9; k++) {
                    double u1 = stream_open(st), u2 = stream_open(st);
                    double r = sqrt(-2.0 * log(u1));
                    m[2 * k] = r * cos(2.0 * M_PI * u2);
                    m[2 * k + 1] = r * sin(2.0 * M_PI * u2);
                }
                or_reunitarize(m);
            }
        }
}

void or_hot_start_haar(double *U, const int n[4], uint64_t seed) { or_hot_start(U, n, seed, OR_HOT_HAAR, 0.0); }

/* One sweep = 8 sub-sweeps (mu = 0..3, parity 0 then 1); a link at site s
 * (parity = x+y+z+t mod 2) draws, per hit, GLAC's 13 centred uniforms for its
 * RST proposal (16-bit halves of seven outputs, stream_rst_uniforms) and one
 * open uniform for the accept test, from its own stream. */
double or_cb_sweep(double *U, const int n[4], double beta, double eps, int nhit,
                   uint64_t seed, uint64_t sweep, int reunit)
{
    long V = volume(n);
    unsigned long long accepted = 0;
    for (int mu = 0; mu < 4; mu++)
        for (int p = 0; p < 2; p++)
            for (long s = 0; s < V; s++) {
                int c[4];
                decode(n, s, c);
                if (((c[0] + c[1] + c[2] + c[3]) & 1) != p) continue;
                uint32_t st[4];
                link_stream(st, seed, s, mu, STREAM_METROPOLIS, sweep);
                double A[18], X[18], Up[18], u[13];
                double *Ul = U + ((long)mu * V + s) * 18;
                staple_at(U, n, c, mu, A);
                for (int h = 0; h < nhit; h++) {
                    stream_rst_uniforms(st, u);
                    or_generate_rst(u, eps, X);
                    or_su3_mul(X, Ul, Up);
                    double dS = or_delta_action(Ul, Up, A, beta);
                    if (exp(-dS) > stream_open(st)) {
                        memcpy(Ul, Up, sizeof Up);
                        accepted++;
                    }
                }
                if (reunit) or_reunitarize_keep_det_sign(Ul);
            }
    return (double)accepted / (4.0 * nhit * (double)V);
}

/* ======================================================================== */
/* Learned proposal (no GLAC counterpart)                                    */
/* ======================================================================== */
/*
 * or_staple_features / or_mlp_forward / or_cholesky_from_output /
 * or_algebra_from_omega / or_project_algebra / or_learned_cb_sweep.
 * Pinned: features are gauge invariant (A -> g A h^dag); the algebra map is
 * i w.T with T = lambda/2 (checked entry by entry against the Gell-Mann
 * matrices) and projection is its adjoint (Re tr(i T_a Y(w)) = w_a / 2
 * orthogonality); the sweep samples exp(beta/3 Re tr(U A)) (single-link
 * histogram) and keeps GLAC's plaquette; or_mlp_forward rounds each layer's
 * input to bf16 to nearest even, as the device's matrix-core GEMMs take it
 * (an identity network separates this from truncation and from no rounding).
 */

static const int STREAM_LEARNED = 3;

void or_staple_features(const double *A, double beta, double f[6])
{
    double Ad[18], B[18], B2[18], d[2];
    or_su3_dag(A, Ad);
    or_su3_mul(A, Ad, B);
    or_su3_mul(B, B, B2);
    or_su3_det(A, d);
    f[0] = B[0] + B[8] + B[16];
    f[1] = B2[0] + B2[8] + B2[16];
    f[2] = d[0] * d[0] + d[1] * d[1];
    f[3] = d[0];
    f[4] = d[1];
    f[5] = beta;
}

static inline double lrelu(double v) { return v > 0.0 ? v : 0.01 * v; }

/* v as FP32, then to bf16 with round-to-nearest-even (finite values) */
static double bf16_round(double v)
{
    float x = (float)v;
    uint32_t u;
    memcpy(&u, &x, 4);
    u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
    memcpy(&x, &u, 4);
    return (double)x;
}

/* The device evaluates the network as bf16 GEMMs with FP32 accumulation:
 * each layer's input (standardised features, hidden activations) is rounded
 * to bf16, as the weights are (the caller passes bf16-valued parameters). */
void or_mlp_forward(const double *P, const double *mean, const double *stdv, const double f[6], double o[36])
{
    const double *W1 = P, *b1 = W1 + 32 * 6, *W2 = b1 + 32, *b2 = W2 + 32 * 32, *W3 = b2 + 32, *b3 = W3 + 36 * 32;
    double x[6], h1[32], h2[32];
    for (int k = 0; k < 6; k++) x[k] = bf16_round((f[k] - mean[k]) / stdv[k]);
    for (int i = 0; i < 32; i++) {
        double s = b1[i];
        for (int k = 0; k < 6; k++) s += W1[i * 6 + k] * x[k];
        h1[i] = bf16_round(lrelu(s));
    }
    for (int i = 0; i < 32; i++) {
        double s = b2[i];
        for (int k = 0; k < 32; k++) s += W2[i * 32 + k] * h1[k];
        h2[i] = bf16_round(lrelu(s));
    }
    for (int i = 0; i < 36; i++) {
        double s = b3[i];
        for (int k = 0; k < 32; k++) s += W3[i * 32 + k] * h2[k];
        o[i] = s;
    }
}

void or_cholesky_from_output(const double o[36], double L[64])
{
    for (int k = 0; k < 64; k++) L[k] = 0.0;
    for (int a = 0; a < 8; a++) {
        L[a * 8 + a] = exp(o[a]);
        for (int b = 0; b < a; b++) L[a * 8 + b] = o[8 + a * (a - 1) / 2 + b];
    }
}

/* Y = i sum_a w_a lambda_a / 2 as a 3x3 matrix in GLAC's layout. */
void or_algebra_from_omega(const double w[8], double Y[18])
{
    const double r3 = 1.0 / sqrt(3.0);
    su3_zero(Y);
    IM(Y, 0, 0) = 0.5 * (w[2] + r3 * w[7]);
    IM(Y, 1, 1) = 0.5 * (-w[2] + r3 * w[7]);
    IM(Y, 2, 2) = -r3 * w[7];
    /* (0,1): lambda_1, lambda_2 ; (0,2): lambda_4, lambda_5 ; (1,2): lambda_6, lambda_7 */
    const int pi[3] = {0, 0, 1}, pj[3] = {1, 2, 2}, re[3] = {1, 4, 6}, im[3] = {0, 3, 5};
    for (int k = 0; k < 3; k++) {
        const int i = pi[k], j = pj[k];
        RE(Y, i, j) = 0.5 * w[re[k]];
        IM(Y, i, j) = 0.5 * w[im[k]];
        RE(Y, j, i) = -0.5 * w[re[k]];
        IM(Y, j, i) = 0.5 * w[im[k]];
    }
}

/* g_a = Re tr(i lambda_a / 2 N) */
void or_project_algebra(const double *N, double g[8])
{
    const double r3 = 1.0 / sqrt(3.0);
    g[0] = -0.5 * (IM(N, 0, 1) + IM(N, 1, 0));
    g[1] = 0.5 * (RE(N, 1, 0) - RE(N, 0, 1));
    g[2] = -0.5 * (IM(N, 0, 0) - IM(N, 1, 1));
    g[3] = -0.5 * (IM(N, 0, 2) + IM(N, 2, 0));
    g[4] = 0.5 * (RE(N, 2, 0) - RE(N, 0, 2));
    g[5] = -0.5 * (IM(N, 1, 2) + IM(N, 2, 1));
    g[6] = 0.5 * (RE(N, 2, 1) - RE(N, 1, 2));
    g[7] = -0.5 * r3 * (IM(N, 0, 0) + IM(N, 1, 1) - 2.0 * IM(N, 2, 2));
}

/* two symmetric unit-variance draws from one 32-bit output: each 16-bit half
 * (low half first) uniform on the 2^16 midpoints of [-1/2, 1/2), times
 * sqrt(12) -- exactly symmetric about 0 */
static inline void stream_sym_pair(uint32_t s[4], double *z0, double *z1)
{
    const uint32_t b = or_xoshiro128ss_next(s);
    *z0 = ((double)(b & 0xFFFFu) * 0x1p-16 - (0.5 - 0x1p-17)) * 3.4641016151377546;
    *z1 = ((double)(b >> 16) * 0x1p-16 - (0.5 - 0x1p-17)) * 3.4641016151377546;
}

double or_learned_cb_sweep(double *U, const int n[4], double beta, double eps, int nhit, uint64_t seed, uint64_t sweep,
                           int reunit, const double *params, const double *mean, const double *stdv)
{
    long V = volume(n);
    unsigned long long accepted = 0;
    for (int mu = 0; mu < 4; mu++)
        for (int p = 0; p < 2; p++)
            for (long s = 0; s < V; s++) {
                int c[4];
                decode(n, s, c);
                if (((c[0] + c[1] + c[2] + c[3]) & 1) != p) continue;
                double A[18], f[6], o[36], L[64], X[18], Y[18], Up[18];
                double *Ul = U + ((long)mu * V + s) * 18;
                staple_at(U, n, c, mu, A);
                or_staple_features(A, beta, f);
                /* staples unlike any in training fall back to the fixed-eps proposal */
                int learned = 1;
                for (int k = 0; k < 6; k++)
                    if (fabs((f[k] - mean[k]) / stdv[k]) > 8.0) learned = 0;
                or_mlp_forward(params, mean, stdv, f, o);
                or_cholesky_from_output(o, L);
                uint32_t st[4];
                link_stream(st, seed, s, mu, STREAM_LEARNED, sweep);
                for (int h = 0; h < nhit; h++) {
                    if (learned) {
                        double z[8], w[8];
                        for (int a = 0; a < 8; a += 2) stream_sym_pair(st, &z[a], &z[a + 1]);
                        for (int a = 0; a < 8; a++) {
                            w[a] = 0.0;
                            for (int b = 0; b <= a; b++) w[a] += L[a * 8 + b] * z[b];
                        }
                        or_algebra_from_omega(w, Y);
                        or_su3_exp(Y, X);
                    } else {
                        double u[13];
                        stream_rst_uniforms(st, u);
                        or_generate_rst(u, eps, X);
                    }
                    or_su3_mul(X, Ul, Up);
                    double dS = or_delta_action(Ul, Up, A, beta);
                    if (exp(-dS) > stream_open(st)) {
                        memcpy(Ul, Up, sizeof Up);
                        accepted++;
                    }
                }
                if (reunit) or_reunitarize_keep_det_sign(Ul);
            }
    return (double)accepted / (4.0 * nhit * (double)V);
}
