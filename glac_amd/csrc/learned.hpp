// learned.hpp -- the staple-conditioned learned Metropolis proposal.
//
// No GLAC counterpart: it replaces the fixed-eps RST proposal
// (GLAC/math/matrices/su3matrixgenerator.h:170-321) when a trained network
// is loaded.  For a link with staple A:
//   features f(A) = (tr AA^dag, tr (AA^dag)^2, |det A|^2, Re det A, Im det A, beta),
//       all invariant under A -> g A h^dagger, standardised with (mean, std);
//   MLP 6 -> 32 -> 32 -> 36 (leaky ReLU, slope 0.01), weights and layer
//       inputs in bf16, FP32 accumulation (matrix cores, mlp_forward_mfma);
//   output o -> lower-triangular L: L_aa = exp(o_a), L_ab = o[8 + a(a-1)/2 + b];
//   per hit w = L z with z symmetric (z_a = c sqrt 12, c uniform on the 2^16
//       midpoints of [-1/2, 1/2), two per 32-bit output), X = exp(i sum_a w_a lambda_a / 2),
//       U' = X U, accepted with GLAC's exp(-dS) > u.
// L depends on A only and z -> -z is a symmetry of the draw, so the
// proposal density of X equals that of X^dagger: exact accept/reject keeps
// detailed balance with respect to exp(-S_W).  If a standardised feature
// leaves [-kOutOfRange, kOutOfRange] (a staple unlike any seen in training)
// the link falls back to the fixed-eps RST proposal; that choice also
// depends on A only.
#pragma once
#include "expsu3.hpp"
#include "lattice.hpp"

namespace glacx {

constexpr int kMlpIn = 6, kMlpH1 = 32, kMlpH2 = 32, kMlpOut = 36;
// outputs per link that pass through a wavefront's LDS buffer in the matrix-core
// forward pass (the other four move between lanes by shuffles)
constexpr int kMlpObuf = 32;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kMlpH1 * kMlpIn, kOffW2 = kOffB1 + kMlpH1,
              kOffB2 = kOffW2 + kMlpH2 * kMlpH1, kOffW3 = kOffB2 + kMlpH2, kOffB3 = kOffW3 + kMlpOut * kMlpH2,
              kNParam = kOffB3 + kMlpOut;
static_assert(kNParam == 2468, "parameter packing");
constexpr float kOutOfRange = 8.0f;

__device__ __forceinline__ float lrelu(float v) { return v > 0.f ? v : 0.01f * v; }

// features of the staple (FP64 accumulation of the invariants)
template <class T> __device__ __forceinline__ void staple_features(const M3<T> &a, double beta, float f[6])
{
    const M3<double> A = convert<double>(a);
    const M3<double> B = mul_ad(A, A);  // A A^dagger, hermitian
    double trb = B.e[0][0].r + B.e[1][1].r + B.e[2][2].r, trb2 = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) trb2 += B.e[i][j].r * B.e[i][j].r + B.e[i][j].i * B.e[i][j].i;
    // det A by cofactors of row 0
    double dr = 0.0, di = 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        const cpx<double> m1 = cmul(A.e[1][j1], A.e[2][j2]), m2 = cmul(A.e[1][j2], A.e[2][j1]);
        const cpx<double> mnr = {m1.r - m2.r, m1.i - m2.i};
        const cpx<double> t = cmul(A.e[0][j], mnr);
        dr += t.r;
        di += t.i;
    }
    f[0] = (float)trb;
    f[1] = (float)trb2;
    f[2] = (float)(dr * dr + di * di);
    f[3] = (float)dr;
    f[4] = (float)di;
    f[5] = (float)beta;
}

// ---- network forward pass on the matrix cores -------------------------------
// A wavefront evaluates the network for its 64 links as three bf16 GEMMs on
// v_mfma_f32_16x16x16_bf16 (FP32 accumulation): weights are the A operand
// (rows = units), activations the B operand (columns = links), 16 links per
// tile.  Lane l of a 16x16 tile holds A[l % 16][4 (l / 16) + 0..3],
// B[4 (l / 16) + 0..3][l % 16] and D[4 (l / 16) + 0..3][l % 16], so the D
// tile of one layer is, element for element, the B fragment of the next:
// only the input features (one link per lane, moved with lane shuffles) and
// the outputs (back to one link per lane: 32 through LDS, the last four by
// lane shuffles) change layout.  Layer
// inputs are rounded to bf16 (round to nearest even), as the weights are.
using bf16x4 = short __attribute__((ext_vector_type(4)));
using f32x4 = float __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned pack_bf16(float a, float b)
{
    return (unsigned)__builtin_bit_cast(unsigned short, (__bf16)a) |
           ((unsigned)__builtin_bit_cast(unsigned short, (__bf16)b) << 16);
}
__device__ __forceinline__ bf16x4 frag_bf16(unsigned lo, unsigned hi)
{
    using U2 = unsigned __attribute__((ext_vector_type(2)));
    U2 v;
    v.x = lo;
    v.y = hi;
    return __builtin_bit_cast(bf16x4, v);
}
__device__ __forceinline__ bf16x4 lrelu_bf16(const f32x4 &d)
{
    return frag_bf16(pack_bf16(lrelu(d[0]), lrelu(d[1])), pack_bf16(lrelu(d[2]), lrelu(d[3])));
}
// The network's parameters as the GEMMs read them (bf16 bits, zero padded so
// every fragment is one aligned 8-byte LDS read): W1 with rows of 8, W3 and
// b3 with 48 rows.
struct MlpImage {
    unsigned short w1[kMlpH1][8];
    unsigned short b1[kMlpH1];
    unsigned short w2[kMlpH2][kMlpH1];
    unsigned short b2[kMlpH2];
    unsigned short w3[48][kMlpH2];
    unsigned short b3[48];
};

// Fill the image from the packed bf16 parameters (all threads of a block).
__device__ __forceinline__ void mlp_image_fill(MlpImage &im, const unsigned short *__restrict__ wbf16, int tid,
                                               int nthreads)
{
    unsigned short *dst = &im.w1[0][0];
    constexpr int n = (int)(sizeof(MlpImage) / sizeof(unsigned short));
    for (int k = tid; k < n; k += nthreads) {
        int src = -1;
        if (k < kMlpH1 * 8) {
            const int r = k >> 3, c = k & 7;
            src = c < kMlpIn ? kOffW1 + r * kMlpIn + c : -1;
        } else if (k < kMlpH1 * 9) {
            src = kOffB1 + (k - kMlpH1 * 8);
        } else if (k < kMlpH1 * 9 + kMlpH2 * kMlpH1) {
            src = kOffW2 + (k - kMlpH1 * 9);
        } else if (k < kMlpH1 * 9 + kMlpH2 * (kMlpH1 + 1)) {
            src = kOffB2 + (k - kMlpH1 * 9 - kMlpH2 * kMlpH1);
        } else if (k < kMlpH1 * 9 + kMlpH2 * (kMlpH1 + 1) + 48 * kMlpH2) {
            const int j = k - (kMlpH1 * 9 + kMlpH2 * (kMlpH1 + 1));
            src = (j >> 5) < kMlpOut ? kOffW3 + j : -1;
        } else {
            const int j = k - (kMlpH1 * 9 + kMlpH2 * (kMlpH1 + 1) + 48 * kMlpH2);
            src = j < kMlpOut ? kOffB3 + j : -1;
        }
        dst[k] = src >= 0 ? wbf16[src] : (unsigned short)0;
    }
}

__device__ __forceinline__ bf16x4 lds_frag(const unsigned short *p)
{
    return __builtin_bit_cast(bf16x4, *reinterpret_cast<const unsigned long long *>(p));
}
__device__ __forceinline__ f32x4 lds_bias(const unsigned short *p)
{
    const unsigned long long v = *reinterpret_cast<const unsigned long long *>(p);
    f32x4 c;
#pragma unroll
    for (int e = 0; e < 4; e++) c[e] = __uint_as_float((unsigned)((v >> (16 * e)) & 0xFFFFu) << 16);
    return c;
}

// Network outputs o[36] of this lane's link; every lane of the wavefront must
// call it (the GEMMs span the wavefront).  im: the parameter image (LDS),
// obuf: this wavefront's 64 x 32 floats of LDS (16-byte aligned).  Returns
// false when a standardised feature is out of range.
__device__ __forceinline__ bool mlp_forward_mfma(const MlpImage &im, const float *mean, const float *stdv,
                                                 const float f[6], float *obuf, float o[36])
{
    const int lane = threadIdx.x & 63, row = lane & 15, grp = lane >> 4;
    float x[kMlpIn];
    bool ok = true;
#pragma unroll
    for (int k = 0; k < kMlpIn; k++) {
        x[k] = (f[k] - mean[k]) / stdv[k];
        ok = ok && fabsf(x[k]) <= kOutOfRange;
    }
    const unsigned p0 = pack_bf16(x[0], x[1]), p1 = pack_bf16(x[2], x[3]), p2 = pack_bf16(x[4], x[5]);
#pragma unroll 1
    for (int nt = 0; nt < 4; nt++) {
        // parameters are re-read from LDS per tile rather than held in registers
        asm volatile("" ::: "memory");
        // features of link 16 nt + l % 16, k = 4 grp + 0..3 (k < 6)
        const int src = 16 * nt + row;
        const unsigned q0 = (unsigned)__shfl((int)p0, src, 64), q1 = (unsigned)__shfl((int)p1, src, 64),
                       q2 = (unsigned)__shfl((int)p2, src, 64);
        const bf16x4 b = frag_bf16(grp == 0 ? q0 : (grp == 1 ? q2 : 0u), grp == 0 ? q1 : 0u);
        bf16x4 a1[2], a2[2];
#pragma unroll
        for (int m = 0; m < 2; m++) {
            bf16x4 w = lds_frag(&im.w1[16 * m + row][4 * (grp & 1)]);
            if (grp >= 2) w = frag_bf16(0u, 0u);
            a1[m] = lrelu_bf16(__builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w, b, lds_bias(&im.b1[16 * m + 4 * grp]), 0, 0, 0));
        }
#pragma unroll
        for (int m = 0; m < 2; m++) {
            f32x4 d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(lds_frag(&im.w2[16 * m + row][4 * grp]), a1[0],
                                                                lds_bias(&im.b2[16 * m + 4 * grp]), 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(lds_frag(&im.w2[16 * m + row][16 + 4 * grp]), a1[1], d, 0, 0, 0);
            a2[m] = lrelu_bf16(d);
        }
#pragma unroll
        for (int m = 0; m < 3; m++) {
            f32x4 d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(lds_frag(&im.w3[16 * m + row][4 * grp]), a2[0],
                                                                lds_bias(&im.b3[16 * m + 4 * grp]), 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(lds_frag(&im.w3[16 * m + row][16 + 4 * grp]), a2[1], d, 0, 0, 0);
            const int r0 = 16 * m + 4 * grp;
            if (m < 2) {
                *reinterpret_cast<f32x4 *>(obuf + (16 * nt + row) * kMlpObuf + r0) = d;
            } else {
                // outputs 32..35 of link 16 nt + row sit in lane row (grp 0):
                // the link's own lane takes them by shuffles
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const float v = __shfl(d[k], lane & 15, 64);
                    if ((lane >> 4) == nt) o[kMlpObuf + k] = v;
                }
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): the wavefront's LDS stores have landed
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int i = 0; i < kMlpObuf; i += 4) {
        const f32x4 v = *reinterpret_cast<const f32x4 *>(obuf + lane * kMlpObuf + i);
        o[i] = v[0];
        o[i + 1] = v[1];
        o[i + 2] = v[2];
        o[i + 3] = v[3];
    }
    return ok;
}

// packed lower triangle, row a holds L[a][0..a] at a(a+1)/2
__device__ __forceinline__ void cholesky_from_output(const float o[36], float L[36])
{
#pragma unroll
    for (int a = 0; a < 8; a++) {
#pragma unroll
        for (int b = 0; b < a; b++) L[a * (a + 1) / 2 + b] = o[8 + a * (a - 1) / 2 + b];
        L[a * (a + 1) / 2 + a] = __expf(o[a]);
    }
}

// Y = i sum_a w_a lambda_a / 2 in compact form
template <class T> __device__ __forceinline__ su3alg<T> algebra_from_omega(const float w[8])
{
    const float r3 = 0.57735026918962576f;
    su3alg<T> y;
    y.a0 = T(0.5f * fmaf(r3, w[7], w[2]));
    y.a1 = T(0.5f * fmaf(r3, w[7], -w[2]));
    y.p[0] = T(0.5f * w[1]);
    y.p[1] = T(0.5f * w[0]);
    y.p[2] = T(0.5f * w[4]);
    y.p[3] = T(0.5f * w[3]);
    y.p[4] = T(0.5f * w[6]);
    y.p[5] = T(0.5f * w[5]);
    return y;
}

// two symmetric unit-variance draws from one 32-bit stream output: each
// 16-bit half (low half first) on the 2^16 midpoints of [-1/2, 1/2) (exact
// in FP32), times sqrt(12).  Bin k and bin 65535 - k give exactly opposite
// values, so z and -z are equally likely.
template <class S> __device__ __forceinline__ void sym_draw_pair(S &rs, float &z0, float &z1)
{
    const uint32_t b = rs.next();
    const float c0 = (float)(b & 0xFFFFu) * 0x1p-16f - (0.5f - 0x1p-17f);
    const float c1 = (float)(b >> 16) * 0x1p-16f - (0.5f - 0x1p-17f);
    z0 = c0 * 3.4641016151377546f;
    z1 = c1 * 3.4641016151377546f;
}

// Re tr(i lambda_a / 2 N) for a = 0..7
template <class T> __device__ __forceinline__ void project_algebra(const M3<T> &n, float g[8])
{
    const float r3 = 0.57735026918962576f;
    g[0] = -0.5f * float(n.e[0][1].i + n.e[1][0].i);
    g[1] = 0.5f * float(n.e[1][0].r - n.e[0][1].r);
    g[2] = -0.5f * float(n.e[0][0].i - n.e[1][1].i);
    g[3] = -0.5f * float(n.e[0][2].i + n.e[2][0].i);
    g[4] = 0.5f * float(n.e[2][0].r - n.e[0][2].r);
    g[5] = -0.5f * float(n.e[1][2].i + n.e[2][1].i);
    g[6] = 0.5f * float(n.e[2][1].r - n.e[1][2].r);
    g[7] = -0.5f * r3 * float(n.e[0][0].i + n.e[1][1].i - 2.f * n.e[2][2].i);
}


} // namespace glacx
