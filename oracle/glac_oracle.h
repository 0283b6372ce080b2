/*
 * glac_oracle.h -- FP64 CPU restatement of GLAC's gauge-generation and
 * gradient-flow path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, as the checker and as the
 * timed CPU baseline.  The product (glac_amd, libglacx.so) never links,
 * loads or calls anything under oracle/.
 *
 * Conventions (GLAC's):
 *  - a link is 18 doubles, row-major 3x3 complex, (re, im) interleaved:
 *    U_ij at m[6i+2j] (re) and m[6i+2j+1] (im)  (GLAC/math/matrices/su3.h:22,64);
 *  - a gauge field is 4 planes, mu-major, each V links in lexicographic site
 *    order x + Nx(y + Ny(z + Nz t))  (GLAC/parallelization/index.h:32-35,
 *    GLAC/system.cpp:158-161);  field offset of U_mu(site) = (mu*V + site)*18;
 *  - n[4] = {Nx, Ny, Nz, Nt}, all periodic.
 *
 * How each function is pinned (the test that would fail on a plausible
 * mistake) is listed in the header of glac_oracle.c.
 */
#ifndef GLAC_ORACLE_H
#define GLAC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Threads the parallel site loops (flow, force, measurement) use; results do
 * not depend on it. */
int or_num_threads(void);

/* ---- 3x3 complex algebra ------------------------------------------------ */
void   or_su3_mul(const double *a, const double *b, double *c);
void   or_su3_dag(const double *a, double *c);
double or_su3_retrace_mul(const double *a, const double *b);
void   or_su3_det(const double *a, double *re_im);

/* ---- random numbers ----------------------------------------------------- */
typedef struct or_mt64 { uint64_t mt[312]; int mti; } or_mt64;
void     or_mt64_seed(or_mt64 *g, uint64_t seed);
uint64_t or_mt64_next(or_mt64 *g);
double   or_mt64_uniform(or_mt64 *g, double a, double b);
void     or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
uint32_t or_xoshiro128ss_next(uint32_t s[4]);

/* ---- GLAC random SU(3) generators, as pure functions of their uniforms -- */
void or_generate_su2(const double r[4], double eps, double out[8]);
void or_rst_mul(const double *r, const double *s, const double *t, double *X);
void or_rst_mul_inv(const double *r, const double *s, const double *t, double *X);
void or_generate_rst(const double u[13], double eps, double X[18]);
void or_generate_random(const double u[18], double H[18]);

/* ---- Wilson action ------------------------------------------------------ */
void   or_staple(const double *U, const int n[4], int x, int y, int z, int t, int mu, double A[18]);
double or_delta_action(const double *Uold, const double *Unew, const double *A, double beta);
void   or_action_derivative(const double *U, const int n[4], int mu, double *Z);

/* ---- SU(3) exponential and the RK3 Wilson flow --------------------------- */
void or_su3_exp(const double *Q, double *out);
void or_su3_exp_taylor(const double *Q, int order, double *out);
void or_flow_step(double *U, const int n[4], double eps);

/* ---- clover observables ------------------------------------------------- */
void or_measure(const double *U, const int n[4], double out[3], double *qt);
void or_measure_density(const double *U, const int n[4], double *qd, double *ed);
void or_measure_super(const double *U, const int n[4], double out[4], double *qt, double *wt);

/* ---- Metropolis: GLAC's sequential sweep with its two mt19937_64 streams -- */
typedef struct or_glac_rng { or_mt64 metro; or_mt64 matrix; } or_glac_rng;
or_glac_rng *or_glac_rng_new(uint64_t metropolis_seed, uint64_t matrix_seed);
void   or_glac_rng_free(or_glac_rng *g);
void   or_glac_hot_start(double *U, const int n[4], or_glac_rng *g, double eps, int rst);
double or_glac_sweep(double *U, const int n[4], or_glac_rng *g, double beta, double eps, int nhit);

/* ---- Metropolis: checkerboard sweep with the counter-based link streams -- */
void   or_reunitarize(double *m);
void   or_reunitarize_keep_det_sign(double *m);
enum { OR_HOT_HAAR = 0, OR_HOT_GLAC_RANDOM = 1, OR_HOT_RST = 2 };
void   or_hot_start(double *U, const int n[4], uint64_t seed, int kind, double eps);
void   or_hot_start_haar(double *U, const int n[4], uint64_t seed);
double or_cb_sweep(double *U, const int n[4], double beta, double eps, int nhit,
                   uint64_t seed, uint64_t sweep, int reunit);

/* ---- learned proposal (no GLAC counterpart) ------------------------------
 * Parameter packing (float64 here, float32 on the device): W1[32][6], b1[32],
 * W2[32][32], b2[32], W3[36][32], b3[36]  (OR_MLP_NPARAM values).
 * Features of a staple A (gauge invariants): tr AA^dag, tr (AA^dag)^2,
 * |det A|^2, Re det A, Im det A, beta; standardised with mean[6], std[6].
 * Output o[36] -> lower-triangular L: L_aa = exp(o_a), L_ab = o[8 + a(a-1)/2 + b].
 * Proposal X = exp(i sum_a w_a lambda_a / 2), w = L z. */
#define OR_MLP_NPARAM 2468
void   or_staple_features(const double *A, double beta, double f[6]);
void   or_mlp_forward(const double *params, const double *mean, const double *stdv, const double f[6],
                      double o[36]);
void   or_cholesky_from_output(const double o[36], double L[64]);
void   or_algebra_from_omega(const double w[8], double Y[18]);
void   or_project_algebra(const double *N, double g[8]);
double or_learned_cb_sweep(double *U, const int n[4], double beta, double eps, int nhit, uint64_t seed,
                           uint64_t sweep, int reunit, const double *params, const double *mean, const double *stdv);

#ifdef __cplusplus
}
#endif
#endif

