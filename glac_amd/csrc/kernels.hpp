// kernels.hpp -- host-side launchers of the gfx950 kernels (one translation
// unit per kernel family) and the shared block reduction.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lattice.hpp"

namespace glacx {

constexpr int kBlock = 256;  // 4 wavefronts of 64

// Flow-stage launch shape: a block covers 64 consecutive sites, wavefront w
// handling direction mu = w, so the four links of a site share a CU.
constexpr int kFlowSites = 64;

struct MetroParams {
    double beta;
    double eps;
    int n_hit;
    int reunit;
    uint64_t seed;
    uint64_t sweep;
};

// Sum `N` per-thread doubles over a block of WAVES wavefronts; thread 0 gets
// the sums.  Wavefront shuffle, then one LDS exchange between the wavefronts;
// the order of additions is fixed, so results are reproducible run to run.
template <int N, int WAVES = kBlock / 64> __device__ __forceinline__ void block_sum(double (&v)[N])
{
    __shared__ double lds[N][WAVES];
#pragma unroll
    for (int k = 0; k < N; k++) {
        double x = v[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        v[k] = x;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < N; k++) lds[k][wave] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; k++) {
            double s = lds[k][0];
#pragma unroll
            for (int w = 1; w < WAVES; w++) s += lds[k][w];
            v[k] = s;
        }
    }
}

// ---- launchers (explicitly instantiated for float and double) ------------
template <class T, int R> void launch_pack(const double *stage, void *f, const Geom &g, int site_major, hipStream_t s);
template <class T, int R> void launch_unpack(const void *f, double *stage, const Geom &g, int site_major, hipStream_t s);
// Hot-start kinds (the values of the C ABI's GLACX_HOT_*).
enum HotKind { HOT_HAAR = 0, HOT_GLAC_RANDOM = 1, HOT_RST = 2 };
template <class T, int R> void launch_hot_start(void *f, const Geom &g, uint64_t seed, int kind, double eps, hipStream_t s);
template <class T, int R> void launch_cold_start(void *f, const Geom &g, hipStream_t s);
template <class T, int R>
void launch_action_derivative(const void *f, double *stage, const Geom &g, int mu, hipStream_t s);

// Clover measurement partials: kMeasAcc doubles {plaq, E, q, w} per
// (chunk, t) slot, chunk c of time slice t at partials[kMeasAcc (c Nt + t)]
// (w, the Weinberg density, only in the SuperSampler variant).
constexpr int kMeasAcc = 4;

// Clover measurement: per-block partial sums, then a one-block fixed-order
// reduction into out = {plaq, E_raw, Q, Q(t)...}, or with `super` (GLAC's
// SuperSampler) {plaq, E_raw, Q, W, Q(t)..., W(t)...}.  mode: 0 plain,
// 1 SuperSampler (Weinberg), 2 plain plus per-site densities into
// density[0..V) (charge) and density[V..2V) (action), GLAC site order.
int measure_partials_count(const Geom &g);
template <class T, int R>
void launch_measure(const void *f, const Geom &g, double *partials, int mode, double *density, hipStream_t s);
void launch_measure_finish(double *partials, const Geom &g, double *out, bool super, hipStream_t s);

int plaquette_partials_count(const Geom &g);
template <class T, int R> void launch_plaquette(const void *f, const Geom &g, double *partials, hipStream_t s);
void launch_plaquette_finish(const double *partials, int nparts, const Geom &g, double *out, hipStream_t s);

// One RK3 stage (0, 1, 2): win -> wout, accumulator read/written in acc.
// With `meas` (stage 0 only, FP32 links, flow_measure_fusable(g)) the stage
// also writes the clover-measurement partials of win into meas, in the layout
// launch_measure_finish reads.
template <class T, int R>
void launch_flow_stage(const void *win, void *wout, void *acc, const Geom &g, int stage, double eps, double *meas,
                       hipStream_t s);
bool flow_measure_fusable(const Geom &g);

// One checkerboard sub-sweep (mu, parity), in place; accepted hits are added
// to *count.
template <class T, int R>
void launch_metropolis(void *f, const Geom &g, int mu, int parity, const MetroParams &p,
                       unsigned long long *count, hipStream_t s);

// ---- learned proposal (k_learned.hip) ----------------------------------------
// Samples are stored column-major: column c (0..35: A then U, 18 reals each)
// of sample s at out[c * cap + s].
template <class T, int R>
void launch_collect(const void *f, const Geom &g, int mu, int parity, int stride, float *out, size_t cap, size_t base,
                    hipStream_t s);
int feature_moments_blocks(size_t n);
void launch_feature_moments(const float *smp, size_t cap, size_t n, double beta, double *partials, float *mean_std,
                            hipStream_t s);
// One training step over `batch` samples (drawn from the first n_samples, or
// the first `batch` in order with draws read from zin when zin != nullptr);
// adam_t > 0 applies an Adam update (bias corrections for step adam_t).
void launch_train_step(const float *smp, size_t cap, size_t n_samples, int batch, int draws, uint64_t seed,
                       uint64_t step, double beta, float *params, const float *mean_std, const float *zin,
                       float *inter, double *loss_partials, float *grad, float *m, float *v, float lr, int adam_t,
                       unsigned short *bf16, hipStream_t s);
void launch_to_bf16(const float *params, unsigned short *bf16, hipStream_t s);
template <class T, int R>
void launch_metropolis_learned(void *f, const Geom &g, int mu, int parity, const MetroParams &p,
                               const unsigned short *wbf16, const float *mean_std, unsigned long long *count,
                               hipStream_t s);
void launch_proposal_eval(const float *feat, int n, const unsigned short *wbf16, const float *mean_std, float *out,
                          int *ok, hipStream_t s);
constexpr int kTrainInterCols = 170;  // per-sample columns kept by the training step
constexpr int kProposalParams = 2468;

} // namespace glacx
