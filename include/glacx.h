/*
 * glacx.h -- C ABI of libglacx.so, the MI355X (gfx950) implementation of
 * GLAC's gauge-generation and gradient-flow hot path.
 *
 * Plain C: pointers, sizes and integer status codes; no C++ or torch types.
 * Every call returns 0 on success and a negative GLACX_E_* code on failure;
 * glacx_last_error() then returns the message.  A context owns one HIP stream
 * (default priority) and all device memory; calls on one context must come
 * from one host thread at a time, and different contexts may be driven from
 * different threads at once (their streams overlap on the GPU, so several
 * small lattices fill it together).  Results written through host pointers
 * are complete when the call returns.
 *
 * Host-side arrays use GLAC's layout: a link is 18 doubles, row-major 3x3
 * complex with (re, im) interleaved (GLAC/math/matrices/su3.h:22,64); a gauge
 * field is four arrays, one per mu, of V links in site order
 * x + Nx(y + Ny(z + Nz t)) (GLAC/parallelization/index.h:32-35,
 * GLAC/system.cpp:158-161).
 *
 * A context holds two gauge fields, GLACX_FIELD_MAIN (GLAC's m_lattice, the
 * Markov chain) and GLACX_FIELD_FLOW (GLAC's m_flowLattice), resident on the
 * device between calls (GLAC/system.cpp:138-188).
 */
#ifndef GLACX_H
#define GLACX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct glacx_ctx glacx_ctx;

enum glacx_precision {
    GLACX_FP64 = 0,           /* links, staples, exponential and sums in FP64 */
    GLACX_FP32_LINKS = 1,     /* FP32 links/staples/flow; FP64 exp coefficients,
                                 reunitarisation and lattice-wide sums */
    GLACX_FP32_COMPRESSED = 2 /* as GLACX_FP32_LINKS, but only the first two rows
                                 of each link are stored; the third is rebuilt
                                 as conj(row0 x row1).  SU(3) links only:
                                 uploads and loads of other links fail */
};

enum glacx_field { GLACX_FIELD_MAIN = 0, GLACX_FIELD_FLOW = 1 };

enum glacx_status {
    GLACX_OK = 0,
    GLACX_E_ARG = -1,         /* bad argument (sizes, null pointers, odd extents) */
    GLACX_E_HIP = -2,         /* HIP runtime error */
    GLACX_E_IO = -3,          /* file missing / short read / write failure */
    GLACX_E_STATE = -4        /* call not valid in the current state */
};

/* Timed kernel families for glacx_timing_read. */
enum glacx_kernel {
    GLACX_K_FLOW_STAGE = 0,
    GLACX_K_MEASURE = 1,
    GLACX_K_METROPOLIS = 2,
    GLACX_K_PLAQUETTE = 3,
    GLACX_K_METROPOLIS_LEARNED = 4,
    GLACX_K_TRAIN = 5,
    GLACX_K_FLOW_STAGE_MEASURE = 6, /* stage 0 with the clover observables fused (FP32 flow with measurement) */
    GLACX_K_COUNT = 7
};

/* Create a context for an Nx x Ny x Nz x Nt periodic lattice (all extents
 * even and >= 2) on HIP device `device`.  Replaces GLAC's lattice allocation
 * in System::subLatticeSetup (GLAC/system.cpp:148-177) and Flow's
 * temporaries (GLAC/flow/flow.cpp:14-25). */
int glacx_create(const int dims[4], int precision, int device, glacx_ctx **out);
int glacx_destroy(glacx_ctx *ctx);
const char *glacx_last_error(const glacx_ctx *ctx);
/* Device bytes held by the context. */
int glacx_device_bytes(const glacx_ctx *ctx, size_t *bytes);

/* Host <-> device field transfer in GLAC's layout: links[mu] points at V*18
 * doubles (a Lattice<SU3>'s site vector).  Replaces the host-side lattice
 * ownership of GLAC/system.cpp:158-161. */
int glacx_upload_field(glacx_ctx *ctx, int field, const double *const links[4]);
int glacx_download_field(glacx_ctx *ctx, int field, double *const links[4]);
/* copyToFlowLattice (GLAC/system.cpp:179-188): dst <- src on the device. */
int glacx_copy_field(glacx_ctx *ctx, int dst, int src);

/* GLAC .bin configuration (GLAC/io/fieldio.cpp:61-122, 195-248): sites in
 * x-fastest order, 4 links per site (mu = 0..3), 18 native-endian doubles
 * per link, no header; 576 V bytes. */
int glacx_load_bin(glacx_ctx *ctx, int field, const char *path);
int glacx_save_bin(glacx_ctx *ctx, int field, const char *path);
/* Chroma configuration (IO::FieldIO::loadChromaFieldConfiguration,
 * GLAC/io/fieldio.cpp:259-338): the same site and link order as .bin, every
 * double big-endian. */
int glacx_load_chroma(glacx_ctx *ctx, int field, const char *path);

/* Starts (System::latticeSetup, GLAC/system.cpp:200-224).  Cold start sets
 * every link to unity.  Hot start draws every link from its own
 * counter-based stream (seed, site, mu), of one of three kinds:
 *   GLACX_HOT_HAAR         Haar-random SU(3) (eps unused);
 *   GLACX_HOT_GLAC_RANDOM  GLAC's default "fully random" start,
 *                          SU3MatrixGenerator::generateRandom
 *                          (GLAC/math/matrices/su3matrixgenerator.h:98-164):
 *                          entries uniform in (-eps, eps), columns
 *                          orthonormalised, det U = -1.  Not available with
 *                          GLACX_FP32_COMPRESSED (returns GLACX_E_ARG);
 *   GLACX_HOT_RST          GLAC's RSTHotStart, generateRST with step eps
 *                          (su3matrixgenerator.h:170-196): SU(3) near unity.
 * eps must lie in (0, 1) for the last two. */
enum glacx_hot_kind { GLACX_HOT_HAAR = 0, GLACX_HOT_GLAC_RANDOM = 1, GLACX_HOT_RST = 2 };
int glacx_cold_start(glacx_ctx *ctx, int field);
int glacx_hot_start(glacx_ctx *ctx, int field, uint64_t seed, int kind, double eps);

/* n_sweeps Metropolis sweeps on GLACX_FIELD_MAIN (System::update,
 * GLAC/system.cpp:333-361, with WilsonGaugeAction::computeStaple /
 * getDeltaAction, GLAC/actions/wilsongaugeaction.cpp:24-53, and GLAC's RST
 * proposal, GLAC/math/matrices/su3matrixgenerator.h:170-321).  A sweep is 8
 * checkerboard sub-sweeps (mu = 0..3, parity even then odd); each link gets
 * n_hit proposals X U accepted iff exp(-dS) > u.  Links are reunitarised
 * (FP64) after their hits when reunit_every > 0 and the global sweep index
 * (sweep_offset + i) is a multiple of reunit_every.  Sweep i draws from the
 * streams of (seed, sweep_offset + i).  accept[i] (may be NULL) receives the
 * acceptance rate accepted / (4 n_hit V) of sweep i (GLAC/system.cpp:359);
 * plaq[i] (may be NULL) the plaquette after sweep i. */
int glacx_metropolis(glacx_ctx *ctx, int n_sweeps, double beta, double eps, int n_hit,
                     int reunit_every, uint64_t seed, uint64_t sweep_offset,
                     double *accept, double *plaq);

/* Clover observables of one field (MasterSamplerTopcXYZ::calculate,
 * GLAC/observables/mastersamplertopcxyz.cpp:236-384): plaquette
 * (normalised to 1 at unity), E_raw (physical E = -E_raw / 64), Q, and
 * q_t[Nt], the charge summed over each time slice.  Any pointer may be NULL. */
int glacx_measure(glacx_ctx *ctx, int field, double *plaq, double *e_raw, double *q, double *q_t);

/* glacx_measure plus the Weinberg operator, as GLAC's SuperSampler
 * (GLAC/observables/supersampler.cpp:298-496): w = W and w_t[Nt] its
 * per-time-slice sums, normalised by GLAC's 2 * 8 * 8 / (3 * 16^3).  Any
 * pointer may be NULL. */
int glacx_measure_super(glacx_ctx *ctx, int field, double *plaq, double *e_raw, double *q, double *w, double *q_t,
                        double *w_t);

/* Per-site charge and action densities of GLAC's LatticeActionChargeDensity
 * (GLAC/observables/latticeactionchargedensity.cpp:169-306), each V doubles in
 * GLAC's site order x + Nx(y + Ny(z + Nz t)): q_density(x) = -sum_nu
 * Re tr(G_{0 nu} G_{rho sigma}) / (256 pi^2) (sums to Q) and e_density(x) =
 * sum_nu [Re tr(G_{0 nu}^2) + Re tr(G_{rho sigma}^2)] (sums to V E_raw), the
 * fields GLAC writes with IO::FieldIO::writeDoublesFieldToFile
 * (GLAC/io/fieldio.cpp:137-186).  Also returns plaq, e_raw, q.  Any pointer
 * may be NULL. */
int glacx_measure_density(glacx_ctx *ctx, int field, double *q_density, double *e_density, double *plaq,
                          double *e_raw, double *q);

/* Plaquette only (Plaquette::calculate, GLAC/observables/plaquette.cpp:26-40). */
int glacx_plaquette(glacx_ctx *ctx, int field, double *plaq);

/* n_steps RK3 Wilson-flow steps of size eps on `field`, in place
 * (Flow::flowField, GLAC/flow/flow.cpp:43-73, with the SU3Exp exponential,
 * GLAC/math/exponentiation/su3exp.cpp:16-79).  When the output pointers are
 * non-NULL the clover observables are measured before the first step and
 * after every step, as System::flowConfiguration does
 * (GLAC/system.cpp:471-517): plaq, e_raw and q hold n_steps + 1 values and
 * q_t holds (n_steps + 1) * Nt values, row i = flow time i * eps.  When
 * b_z Ny Nx / 2 is a multiple of 32 (b_z = 4 if 4 divides Nz, else 2; so
 * Nx Ny divisible by 16, or by 32 when 4 does not divide Nz), each step's
 * first RK3 stage measures the field it reads, so a measured flow costs
 * little more than an unmeasured one. */
int glacx_flow(glacx_ctx *ctx, int field, int n_steps, double eps,
               double *plaq, double *e_raw, double *q, double *q_t);

/* glacx_flow measuring with the SuperSampler observables (GLAC's flow with
 * "weinberg" in the observable list, GLAC/system.cpp:76-90): w holds n_steps + 1
 * values and w_t (n_steps + 1) * Nt, laid out as q and q_t. */
int glacx_flow_super(glacx_ctx *ctx, int field, int n_steps, double eps, double *plaq, double *e_raw, double *q,
                     double *w, double *q_t, double *w_t);

/* glacx_flow that allocates nothing.  No counterpart in GLAC: for callers
 * that cannot allocate device memory during the call (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * measurement partials and per-step result rows live in a fixed 16 MiB
 * device array, one per GPU, held until the process ends; the call returns
 * GLACX_E_ARG when (n_steps + 1) rows of (3 + Nt) doubles plus the clover
 * reduction's partials (4 Nt (V / (Nt 64) + 1) doubles, about 0.5 MiB at
 * 32^4) exceed it -- split such a flow into shorter calls.  Calls that
 * overlap on one GPU (several contexts or threads) share the array and race;
 * serialise them.  Event timing (glacx_timing_enable) does not record these
 * calls.  Arguments and results as glacx_flow. */
int glacx_flow_noalloc(glacx_ctx *ctx, int field, int n_steps, double eps,
                       double *plaq, double *e_raw, double *q, double *q_t);

/* Whole-lattice flow force for one direction in GLAC's layout, V*18 doubles
 * (WilsonGaugeAction::getActionDerivative,
 * GLAC/actions/wilsongaugeaction.cpp:55-86). */
int glacx_action_derivative(glacx_ctx *ctx, int field, int mu, double *z_out);

/* ---- Learned proposal.  No counterpart in GLAC: it replaces the fixed-eps
 * RST proposal (GLAC/math/matrices/su3matrixgenerator.h:170-321) in the
 * Metropolis update (GLAC/system.cpp:321-361).  A network maps six gauge
 * invariants of the staple to the Cholesky factor of the covariance of an
 * su(3) step; parameters are 2468 floats packed W1[32][6], b1[32],
 * W2[32][32], b2[32], W3[36][32], b3[36]; features are standardised with
 * mean[6], std[6]. */

/* Start a new training-sample buffer holding up to `capacity` samples. */
int glacx_samples_reset(glacx_ctx *ctx, size_t capacity);
/* Append (staple, link) pairs of every `stride`-th link of each direction
 * and parity of GLACX_FIELD_MAIN; *count receives the total held. */
int glacx_samples_collect(glacx_ctx *ctx, int stride, size_t *count);
/* Feature mean and standard deviation over the held samples; also stored
 * as the context's standardisation. */
int glacx_samples_stats(glacx_ctx *ctx, double beta, float mean[6], float std[6]);
/* Set / get the network parameters and standardisation. */
int glacx_proposal_set(glacx_ctx *ctx, const float *params, const float mean[6], const float std[6]);
int glacx_proposal_get(glacx_ctx *ctx, float *params, float mean[6], float std[6]);
/* `steps` Adam steps (learning rate lr) on batches of `batch` samples with
 * `draws` proposal draws each; loss[i] (may be NULL) receives each step's
 * loss (the negative expected squared jump distance). */
int glacx_train(glacx_ctx *ctx, int steps, int batch, int draws, double lr, double beta, uint64_t seed,
                double *loss);
/* Loss and gradient of the current parameters on n host samples (n x 36
 * floats: staple then link, GLAC's 18-real layout each) with the given
 * draws z (n x draws x 8 floats): the device training step without the
 * update, for checking. */
int glacx_train_grad(glacx_ctx *ctx, const float *samples, int n, const float *z, int draws, double beta,
                     double *loss, float *grad);
/* The network as the learned sweeps evaluate it (bf16 matrix-core GEMMs,
 * FP32 accumulation) on n host rows of the six staple features: 36 outputs
 * per row into out (n x 36 floats), and in_range[i] (may be NULL) = 1 when
 * row i's standardised features lie in [-8, 8] (else the sweep falls back to
 * the fixed-eps proposal).  For checking, like glacx_train_grad. */
int glacx_proposal_eval(glacx_ctx *ctx, const float *features, int n, float *out, int *in_range);
/* glacx_metropolis with the learned proposal (fixed-eps RST, step eps, for
 * links whose standardised features fall outside [-8, 8]). */
int glacx_metropolis_learned(glacx_ctx *ctx, int n_sweeps, double beta, double eps, int n_hit, int reunit_every,
                             uint64_t seed, uint64_t sweep_offset, double *accept, double *plaq);

/* Block until the context's stream is idle. */
int glacx_synchronize(glacx_ctx *ctx);

/* Per-kernel device time with HIP events on the context's stream.  While
 * enabled, every launch of a timed kernel family is bracketed by events;
 * glacx_timing_read synchronises and returns the launch count and the summed
 * device milliseconds since the last reset. */
int glacx_timing_enable(glacx_ctx *ctx, int on);
int glacx_timing_reset(glacx_ctx *ctx);
int glacx_timing_read(glacx_ctx *ctx, int kernel, long long *launches, double *total_ms);

#ifdef __cplusplus
}
#endif
#endif /* GLACX_H */
