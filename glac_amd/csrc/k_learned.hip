// k_learned.hip -- training and sampling kernels of the learned proposal
// (definitions in learned.hpp; no GLAC counterpart).
//
// Training data: (staple A, link U) pairs read off the chain of the plain
// sampler.  Objective, per sample and per draw z (W = U A, w = L(A) z,
// Y = i w.lambda / 2):
//     dS_2 = -(beta / 3) (Re tr(Y W) + Re tr(Y Y W) / 2),  alpha = min(1, e^{-dS_2}),
//     loss = -alpha |w|^2              (negative expected squared jump distance).
// The second-order dS only shapes the proposal; the sampler uses the exact
// exponential and exact dS.
//
// A training step is three kernels: train_kernel (one thread per sample:
// forward, loss, per-sample back-propagation to the layer inputs, written as
// columns), grad_kernel (one thread per parameter: the batch sum of the
// outer products, in a fixed order) and adam_kernel.
#include "kernels.hpp"
#include "learned.hpp"
#include "proposal.hpp"

namespace glacx {

// Samples hold 36 columns per sample: A (18 reals) then U (18 reals).
// Per-sample columns kept for the gradient: x, h1, h2, do, da2, da1.
constexpr int kColX = 0, kColH1 = kColX + kMlpIn, kColH2 = kColH1 + kMlpH1, kColDo = kColH2 + kMlpH2,
              kColDa2 = kColDo + kMlpOut, kColDa1 = kColDa2 + kMlpH2;
static_assert(kColDa1 + kMlpH1 == kTrainInterCols, "intermediate column count");

template <class T> __device__ __forceinline__ M3<float> to_float(const M3<T> &m) { return convert<float>(m); }

__device__ __forceinline__ M3<float> read_matrix(const float *__restrict__ s, size_t cap, int col0, size_t i)
{
    M3<float> m;
#pragma unroll
    for (int k = 0; k < 9; k++) m.e[k / 3][k % 3] = {s[(size_t)(col0 + 2 * k) * cap + i], s[(size_t)(col0 + 2 * k + 1) * cap + i]};
    return m;
}

// ---- sample collection -----------------------------------------------------
template <class T, int R>
__global__ __launch_bounds__(kBlock) void collect_kernel(const typename vec2<T>::type *__restrict__ f, Geom g, int mu,
                                                         int parity, int stride, float *__restrict__ out, size_t cap,
                                                         size_t base)
{
    const int h = (blockIdx.x * kBlock + threadIdx.x) * stride;
    if (h >= g.vh) return;
    const size_t s = base + (size_t)(h / stride);
    if (s >= cap) return;
    const int idx = parity_index(g, parity, h);
    const Site x = site_of(g, idx);
    const M3<T> a = staple<T, R>(f, g, x, mu);
    const M3<T> u = load_link<T, R>(f, g, mu, idx);
#pragma unroll
    for (int k = 0; k < 9; k++) {
        out[(size_t)(2 * k) * cap + s] = float(a.e[k / 3][k % 3].r);
        out[(size_t)(2 * k + 1) * cap + s] = float(a.e[k / 3][k % 3].i);
        out[(size_t)(18 + 2 * k) * cap + s] = float(u.e[k / 3][k % 3].r);
        out[(size_t)(18 + 2 * k + 1) * cap + s] = float(u.e[k / 3][k % 3].i);
    }
}

// ---- feature standardisation ---------------------------------------------------
__global__ __launch_bounds__(kBlock) void feature_moments_kernel(const float *__restrict__ smp, size_t cap, size_t n,
                                                                 double beta, double *__restrict__ partials)
{
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    double acc[12];
#pragma unroll
    for (int k = 0; k < 12; k++) acc[k] = 0.0;
    if (i < n) {
        float fe[6];
        staple_features(read_matrix(smp, cap, 0, i), beta, fe);
#pragma unroll
        for (int k = 0; k < 6; k++) {
            acc[k] = fe[k];
            acc[6 + k] = (double)fe[k] * fe[k];
        }
    }
    block_sum<12>(acc);
    if (threadIdx.x == 0)
        for (int k = 0; k < 12; k++) partials[(size_t)blockIdx.x * 12 + k] = acc[k];
}

__global__ void feature_moments_finish(const double *__restrict__ partials, int nblk, size_t n,
                                       float *__restrict__ mean_std)
{
    if (threadIdx.x != 0) return;
    for (int k = 0; k < 6; k++) {
        double s = 0.0, s2 = 0.0;
        for (int b = 0; b < nblk; b++) {
            s += partials[(size_t)b * 12 + k];
            s2 += partials[(size_t)b * 12 + 6 + k];
        }
        const double m = s / (double)n;
        const double var = s2 / (double)n - m * m;
        mean_std[k] = (float)m;
        mean_std[6 + k] = (float)(var > 1e-12 ? sqrt(var) : 1.0);
    }
}

// ---- training step -------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void train_kernel(const float *__restrict__ smp, size_t cap, size_t n_samples,
                                                       int batch, int draws, uint64_t seed, uint64_t step,
                                                       double beta, const float *__restrict__ params,
                                                       const float *__restrict__ mean_std,
                                                       const float *__restrict__ zin, float *__restrict__ inter,
                                                       double *__restrict__ loss_partials)
{
    __shared__ float P[kNParam];
    for (int k = threadIdx.x; k < kNParam; k += kBlock) P[k] = params[k];
    __syncthreads();
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double lacc[1] = {0.0};
    if (i < batch) {
        size_t s;
        if (zin) {
            s = (size_t)i;
        } else {
            uint32_t c[4] = {(uint32_t)i, (uint32_t)(step >> 32), (uint32_t)step, 0x7E57u};
            philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
            s = (size_t)((((uint64_t)c[0] << 32) | c[1]) % n_samples);
        }
        const M3<float> A = read_matrix(smp, cap, 0, s);
        const M3<float> U = read_matrix(smp, cap, 18, s);
        float fe[6];
        staple_features(A, beta, fe);
        // forward; the layer inputs go straight to their columns and only the
        // signs of the pre-activations are kept for the backward pass
        float *col = inter + i;
        float x[kMlpIn], h1[kMlpH1], o[kMlpOut];
        uint32_t pos1 = 0u, pos2 = 0u;
#pragma unroll
        for (int k = 0; k < kMlpIn; k++) {
            x[k] = (fe[k] - mean_std[k]) / mean_std[6 + k];
            col[(size_t)(kColX + k) * batch] = x[k];
        }
#pragma unroll
        for (int r = 0; r < kMlpH1; r++) {
            float v = P[kOffB1 + r];
#pragma unroll
            for (int k = 0; k < kMlpIn; k++) v = fmaf(P[kOffW1 + r * kMlpIn + k], x[k], v);
            pos1 |= (v > 0.f ? 1u : 0u) << r;
            h1[r] = lrelu(v);
            col[(size_t)(kColH1 + r) * batch] = h1[r];
        }
        {
            float h2[kMlpH2];
#pragma unroll
            for (int r = 0; r < kMlpH2; r++) {
                float v = P[kOffB2 + r];
#pragma unroll
                for (int k = 0; k < kMlpH1; k++) v = fmaf(P[kOffW2 + r * kMlpH1 + k], h1[k], v);
                pos2 |= (v > 0.f ? 1u : 0u) << r;
                h2[r] = lrelu(v);
                col[(size_t)(kColH2 + r) * batch] = h2[r];
            }
#pragma unroll
            for (int r = 0; r < kMlpOut; r++) {
                float v = P[kOffB3 + r];
#pragma unroll
                for (int k = 0; k < kMlpH2; k++) v = fmaf(P[kOffW3 + r * kMlpH2 + k], h2[k], v);
                o[r] = v;
            }
        }
        float L[36];
        cholesky_from_output(o, L);
        const M3<float> W = mul(U, A);
        float gW[8];
        project_algebra(W, gW);
        float dL[36];
#pragma unroll
        for (int k = 0; k < 36; k++) dL[k] = 0.f;
        const float bf = (float)(beta / 3.0);
        const float norm = 1.0f / ((float)batch * (float)draws);
        Stream rs(seed, (uint32_t)i, 0u, STREAM_TRAIN, step);
        double loss = 0.0;
#pragma unroll 1
        for (int k = 0; k < draws; k++) {
            float z[8], w[8];
            if (zin) {
#pragma unroll
                for (int a = 0; a < 8; a++) z[a] = zin[((size_t)i * draws + k) * 8 + a];
            } else {
#pragma unroll
                for (int a = 0; a < 8; a += 2) sym_draw_pair(rs, z[a], z[a + 1]);
            }
            float w2 = 0.f;
#pragma unroll
            for (int a = 0; a < 8; a++) {
                float v = 0.f;
#pragma unroll
                for (int b = 0; b <= a; b++) v = fmaf(L[a * (a + 1) / 2 + b], z[b], v);
                w[a] = v;
                w2 = fmaf(v, v, w2);
            }
            const M3<float> Y = expand(algebra_from_omega<float>(w));
            const M3<float> YW = mul(Y, W);
            M3<float> Q = mul(W, Y);  // Q = YW + WY
            add_to(Q, YW);
            const float lin = retrace(YW);
            const float quad = retrace_mul(Y, YW);
            const float dS = -bf * (lin + 0.5f * quad);
            const float alpha = dS > 0.f ? __expf(-dS) : 1.f;
            loss -= (double)(alpha * w2) * norm;
            float gQ[8];
            project_algebra(Q, gQ);
            float dw[8];
#pragma unroll
            for (int a = 0; a < 8; a++) {
                const float ddS = -bf * (gW[a] + 0.5f * gQ[a]);
                const float dalpha = dS > 0.f ? -alpha * ddS : 0.f;
                dw[a] = -(2.f * alpha * w[a] + w2 * dalpha) * norm;
            }
#pragma unroll
            for (int a = 0; a < 8; a++)
#pragma unroll
                for (int b = 0; b <= a; b++) dL[a * (a + 1) / 2 + b] = fmaf(dw[a], z[b], dL[a * (a + 1) / 2 + b]);
        }
        lacc[0] = loss;
        // d loss / d o
        float dout[kMlpOut];
#pragma unroll
        for (int a = 0; a < 8; a++) {
            dout[a] = dL[a * (a + 1) / 2 + a] * L[a * (a + 1) / 2 + a];
#pragma unroll
            for (int b = 0; b < a; b++) dout[8 + a * (a - 1) / 2 + b] = dL[a * (a + 1) / 2 + b];
        }
        float da2[kMlpH2];
#pragma unroll
        for (int r = 0; r < kMlpH2; r++) {
            float v = 0.f;
#pragma unroll
            for (int q = 0; q < kMlpOut; q++) v = fmaf(P[kOffW3 + q * kMlpH2 + r], dout[q], v);
            da2[r] = (pos2 >> r) & 1u ? v : 0.01f * v;
            col[(size_t)(kColDa2 + r) * batch] = da2[r];
        }
#pragma unroll
        for (int k = 0; k < kMlpOut; k++) col[(size_t)(kColDo + k) * batch] = dout[k];
#pragma unroll
        for (int r = 0; r < kMlpH1; r++) {
            float v = 0.f;
#pragma unroll
            for (int q = 0; q < kMlpH2; q++) v = fmaf(P[kOffW2 + q * kMlpH1 + r], da2[q], v);
            col[(size_t)(kColDa1 + r) * batch] = (pos1 >> r) & 1u ? v : 0.01f * v;
        }
    }
    block_sum<1>(lacc);
    if (threadIdx.x == 0) loss_partials[blockIdx.x] = lacc[0];
}

// One thread per parameter: sum over the batch, in order, of the outer
// product that gives its gradient.
__global__ __launch_bounds__(kBlock) void grad_kernel(const float *__restrict__ inter, int batch,
                                                      float *__restrict__ grad)
{
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= kNParam) return;
    int ca, cb;  // columns multiplied (cb < 0: bias, sum of ca)
    if (p < kOffB1) {
        ca = kColDa1 + (p - kOffW1) / kMlpIn;
        cb = kColX + (p - kOffW1) % kMlpIn;
    } else if (p < kOffW2) {
        ca = kColDa1 + (p - kOffB1);
        cb = -1;
    } else if (p < kOffB2) {
        ca = kColDa2 + (p - kOffW2) / kMlpH1;
        cb = kColH1 + (p - kOffW2) % kMlpH1;
    } else if (p < kOffW3) {
        ca = kColDa2 + (p - kOffB2);
        cb = -1;
    } else if (p < kOffB3) {
        ca = kColDo + (p - kOffW3) / kMlpH2;
        cb = kColH2 + (p - kOffW3) % kMlpH2;
    } else {
        ca = kColDo + (p - kOffB3);
        cb = -1;
    }
    const float *A = inter + (size_t)ca * batch;
    double s = 0.0;
    if (cb < 0) {
        for (int i = 0; i < batch; i++) s += A[i];
    } else {
        const float *B = inter + (size_t)cb * batch;
        for (int i = 0; i < batch; i++) s += (double)A[i] * B[i];
    }
    grad[p] = (float)s;
}

__device__ __forceinline__ unsigned short float_to_bf16(float v)
{
    const unsigned u = __float_as_uint(v);
    const unsigned r = u + 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (unsigned short)(r >> 16);
}

__global__ __launch_bounds__(kBlock) void adam_kernel(float *__restrict__ params, float *__restrict__ m,
                                                      float *__restrict__ v, const float *__restrict__ grad, float lr,
                                                      float c1, float c2, unsigned short *__restrict__ bf16)
{
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= kNParam) return;
    const float g = grad[p];
    const float mm = 0.9f * m[p] + 0.1f * g;
    const float vv = 0.999f * v[p] + 0.001f * g * g;
    m[p] = mm;
    v[p] = vv;
    const float upd = lr * (mm / c1) / (sqrtf(vv / c2) + 1e-8f);
    params[p] -= upd;
    bf16[p] = float_to_bf16(params[p]);
}

__global__ __launch_bounds__(kBlock) void to_bf16_kernel(const float *__restrict__ params,
                                                         unsigned short *__restrict__ bf16)
{
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p < kNParam) bf16[p] = float_to_bf16(params[p]);
}

// ---- sampling with the learned proposal ----------------------------------------
// Waves per SIMD the FP32 sampling kernel is compiled for (build-time tuning):
// four fit 128 VGPRs and, with the network's output buffer at 32 floats per
// link, 38.7 KB of LDS per block.
#ifndef GLACX_LEARNED_WAVES
#define GLACX_LEARNED_WAVES 4
#endif
template <class T, int R>
__global__ __launch_bounds__(kBlock, sizeof(T) == 4 ? GLACX_LEARNED_WAVES : 1) void metropolis_learned_kernel(typename vec2<T>::type *__restrict__ f, Geom g,
                                                                    int mu, int parity, MetroParams prm,
                                                                    const unsigned short *__restrict__ wbf16,
                                                                    const float *__restrict__ mean_std,
                                                                    unsigned long long *__restrict__ count)
{
    __shared__ __align__(16) MlpImage im;
    __shared__ __align__(16) float obuf[kBlock / 64][64 * kMlpObuf];
    mlp_image_fill(im, wbf16, threadIdx.x, kBlock);
    __syncthreads();
    const int h_raw = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    const bool valid = h_raw < g.vh;
    const int h = valid ? h_raw : g.vh - 1;  // idle lanes mirror a real link for the network GEMMs, store nothing
    double nacc[1] = {0.0};
    {
        const int idx = parity_index(g, parity, h);
        const Site x = site_of(g, idx);
        const M3<T> a = staple<T, R>(f, g, x, mu);
        M3<T> u = load_link<T, R>(f, g, mu, idx);
        float fe[6], o[36], L[36];
        staple_features(a, prm.beta, fe);
        const bool learned =
            mlp_forward_mfma(im, mean_std, mean_std + 6, fe, obuf[threadIdx.x >> 6], o);  // whole wavefront
        cholesky_from_output(o, L);
        Stream rs(prm.seed, (uint32_t)lex_of(g, x), (uint32_t)mu, STREAM_LEARNED, prm.sweep);
        const double bf = prm.beta / 3.0;
        const T eps = T(prm.eps);
        const T x0abs = sqrt(T(1) - eps * eps);
        T tra = retrace_mul(u, a);
        int accepted = 0;
#pragma unroll 1
        for (int hit = 0; hit < prm.n_hit; hit++) {
            M3<T> xm;
            if (learned) {
                float z[8], om[8];
#pragma unroll
                for (int k = 0; k < 8; k += 2) sym_draw_pair(rs, z[k], z[k + 1]);
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    float v = 0.f;
#pragma unroll
                    for (int b = 0; b <= k; b++) v = fmaf(L[k * (k + 1) / 2 + b], z[b], v);
                    om[k] = v;
                }
                xm = exp_su3_proposal(algebra_from_omega<T>(om));
            } else {
                xm = rst_proposal<T>(rs, x0abs, eps);
            }
            const M3<T> v = propose_link<T, R>(xm, u);
            const T trv = retrace_mul(v, a);
            const T ds = T(-bf) * (trv - tra);
            if (accept_prob(ds) > rs.open()) {
                u = v;
                tra = trv;
                accepted++;
            }
        }
        if (valid) {
            if (prm.reunit) {
                M3<double> d = convert<double>(u);
                if constexpr (R == 2)
                    reunitarize<2, true>(d);  // the store keeps rows 0 and 1 only
                else
                    reunitarize_keep_det_sign<sizeof(T) == 4>(d);  // det U = +1 or -1 kept
                u = convert<T>(d);
            }
            store_link<T, R>(f, g, mu, idx, u);
            nacc[0] = (double)accepted;
        }
    }
    block_sum<1>(nacc);
    if (threadIdx.x == 0) atomicAdd(count, (unsigned long long)nacc[0]);
}

// The sampling kernels' matrix-core forward pass on n given feature rows (6
// floats each): 36 outputs per row into out, and whether the row's
// standardised features are in range into ok.  A check of the network the
// sweeps evaluate; idle lanes of the last block mirror row n - 1.
__global__ __launch_bounds__(kBlock) void proposal_eval_kernel(const float *__restrict__ feat, int n,
                                                               const unsigned short *__restrict__ wbf16,
                                                               const float *__restrict__ mean_std,
                                                               float *__restrict__ out, int *__restrict__ ok)
{
    __shared__ __align__(16) MlpImage im;
    __shared__ __align__(16) float obuf[kBlock / 64][64 * kMlpObuf];
    mlp_image_fill(im, wbf16, threadIdx.x, kBlock);
    __syncthreads();
    const int i_raw = blockIdx.x * kBlock + threadIdx.x;
    const int i = i_raw < n ? i_raw : n - 1;
    float fe[kMlpIn], o[kMlpOut];
#pragma unroll
    for (int k = 0; k < kMlpIn; k++) fe[k] = feat[(size_t)kMlpIn * i + k];
    const bool in = mlp_forward_mfma(im, mean_std, mean_std + kMlpIn, fe, obuf[threadIdx.x >> 6], o);
    if (i_raw < n) {
#pragma unroll
        for (int k = 0; k < kMlpOut; k++) out[(size_t)kMlpOut * i + k] = o[k];
        ok[i] = in ? 1 : 0;
    }
}

// ---- launchers -------------------------------------------------------------------
void launch_proposal_eval(const float *feat, int n, const unsigned short *wbf16, const float *mean_std, float *out,
                          int *ok, hipStream_t s)
{
    proposal_eval_kernel<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(feat, n, wbf16, mean_std, out, ok);
}

template <class T, int R>
void launch_collect(const void *f, const Geom &g, int mu, int parity, int stride, float *out, size_t cap, size_t base,
                    hipStream_t s)
{
    const int n = (g.vh + stride - 1) / stride;
    collect_kernel<T, R><<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>((const typename vec2<T>::type *)f, g, mu, parity,
                                                                   stride, out, cap, base);
}

int feature_moments_blocks(size_t n) { return (int)((n + kBlock - 1) / kBlock); }

void launch_feature_moments(const float *smp, size_t cap, size_t n, double beta, double *partials, float *mean_std,
                            hipStream_t s)
{
    const int nb = feature_moments_blocks(n);
    feature_moments_kernel<<<nb, kBlock, 0, s>>>(smp, cap, n, beta, partials);
    feature_moments_finish<<<1, 64, 0, s>>>(partials, nb, n, mean_std);
}

void launch_train_step(const float *smp, size_t cap, size_t n_samples, int batch, int draws, uint64_t seed,
                       uint64_t step, double beta, float *params, const float *mean_std, const float *zin,
                       float *inter, double *loss_partials, float *grad, float *m, float *v, float lr, int adam_t,
                       unsigned short *bf16, hipStream_t s)
{
    train_kernel<<<(batch + kBlock - 1) / kBlock, kBlock, 0, s>>>(smp, cap, n_samples, batch, draws, seed, step, beta,
                                                                 params, mean_std, zin, inter, loss_partials);
    grad_kernel<<<(kNParam + kBlock - 1) / kBlock, kBlock, 0, s>>>(inter, batch, grad);
    if (adam_t > 0) {
        const float c1 = 1.f - powf(0.9f, (float)adam_t), c2 = 1.f - powf(0.999f, (float)adam_t);
        adam_kernel<<<(kNParam + kBlock - 1) / kBlock, kBlock, 0, s>>>(params, m, v, grad, lr, c1, c2, bf16);
    }
}

void launch_to_bf16(const float *params, unsigned short *bf16, hipStream_t s)
{
    to_bf16_kernel<<<(kNParam + kBlock - 1) / kBlock, kBlock, 0, s>>>(params, bf16);
}

template <class T, int R>
void launch_metropolis_learned(void *f, const Geom &g, int mu, int parity, const MetroParams &p,
                               const unsigned short *wbf16, const float *mean_std, unsigned long long *count,
                               hipStream_t s)
{
    const unsigned blocks = (unsigned)((g.vh + kBlock - 1) / kBlock);
    metropolis_learned_kernel<T, R><<<blocks, kBlock, 0, s>>>((typename vec2<T>::type *)f, g, mu, parity, p, wbf16,
                                                           mean_std, count);
}

template void launch_collect<float, 3>(const void *, const Geom &, int, int, int, float *, size_t, size_t, hipStream_t);
template void launch_collect<float, 2>(const void *, const Geom &, int, int, int, float *, size_t, size_t, hipStream_t);
template void launch_collect<double, 3>(const void *, const Geom &, int, int, int, float *, size_t, size_t, hipStream_t);
template void launch_metropolis_learned<float, 3>(void *, const Geom &, int, int, const MetroParams &,
                                               const unsigned short *, const float *, unsigned long long *,
                                               hipStream_t);
template void launch_metropolis_learned<float, 2>(void *, const Geom &, int, int, const MetroParams &,
                                               const unsigned short *, const float *, unsigned long long *,
                                               hipStream_t);
template void launch_metropolis_learned<double, 3>(void *, const Geom &, int, int, const MetroParams &,
                                                const unsigned short *, const float *, unsigned long long *,
                                                hipStream_t);

} // namespace glacx
