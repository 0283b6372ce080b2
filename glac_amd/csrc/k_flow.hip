// k_flow.hip -- one fused stage of Luscher's RK3 Wilson flow.
//
// GLAC's Flow::flowField (GLAC/flow/flow.cpp:43-73) runs each stage as four
// whole-lattice getActionDerivative passes (GLAC/actions/wilsongaugeaction.cpp:
// 55-86), a scale/combine pass, and four matrixExp passes (flow.cpp:95-115),
// each materialising lattice temporaries.  Here a stage is one kernel per
// link: staple -> Omega = U A -> Z = -P_TA(Omega) -> RK3 combination with the
// accumulator -> exp (FP64 Cayley-Hamilton coefficients) -> X U, reading the
// field from `win` and writing `wout` (ping-pong, so every Z of a stage sees
// the same field, as GLAC's separate loops do).
//
//   stage 0:  acc <- Z0                       W1 = exp(eps/4 Z0) W0
//   stage 1:  Y = 8/9 eps Z1 - 17/36 eps acc;  acc <- Y;  W2 = exp(Y) W1
//   stage 2:  Y = 3/4 eps Z2 - acc;                      W3 = exp(Y) W2
//
// The accumulator holds the 8 real parameters of an su(3) element per link
// (16-byte vectors, see acc_store).
//
// FP32 links are reunitarised (FP64 Gram-Schmidt) as they are stored.
//
// Launch: 256-thread blocks = 64 sites x 4 directions; wavefront w handles
// mu = w for the block's 64 sites, the h-th to (h+31)-th of each parity (two
// contiguous runs of plane entries, side by side in the lattice).  Even and
// odd sites are neighbours of each other, so a block reads the links around
// one spot of both parity halves, and the four links of a site -- which
// share most staple links -- are processed on the same CU at the same time.
#include "expsu3.hpp"
#include "kernels.hpp"
#include "ldsmat.hpp"

namespace glacx {

// Links a block shares through LDS: wavefront mu loads U_mu(x) (its own
// link) and U_mu(x - mu) for the block's 64 sites; the other three
// wavefronts take them from here for their staples instead of gathering them
// again (U_nu(x) closes every upper staple, U_nu(x - nu) every lower one).
// Per site that is 56 link gathers from memory instead of 76.
// slots 0..3: U_mu(x) for mu = 0..3, slots 4..7: U_mu(x - mu)
template <class T> using FlowLds = LdsMats<T, 8>;

// RK3 accumulator: the 8 real parameters (a0, a1, p0..p5) of an su(3)
// element per link, in 16-byte vectors: two float4 planes per direction in
// FP32 (a0 a1 p0 p1 | p2 p3 p4 p5), four double2 planes in FP64.
__device__ __forceinline__ void acc_store(float *acc, const Geom &g, int mu, int idx, const su3alg<float> &z)
{
    float4 *q = reinterpret_cast<float4 *>(acc) + (size_t)mu * 2 * g.v + idx;
    q[0] = make_float4(z.a0, z.a1, z.p[0], z.p[1]);
    q[(size_t)g.v] = make_float4(z.p[2], z.p[3], z.p[4], z.p[5]);
}
__device__ __forceinline__ su3alg<float> acc_load(const float *acc, const Geom &g, int mu, int idx)
{
    const float4 *q = reinterpret_cast<const float4 *>(acc) + (size_t)mu * 2 * g.v + idx;
    const float4 u = q[0], v = q[(size_t)g.v];
    su3alg<float> z;
    z.a0 = u.x, z.a1 = u.y, z.p[0] = u.z, z.p[1] = u.w, z.p[2] = v.x, z.p[3] = v.y, z.p[4] = v.z, z.p[5] = v.w;
    return z;
}
__device__ __forceinline__ void acc_store(double *acc, const Geom &g, int mu, int idx, const su3alg<double> &z)
{
    double2 *q = reinterpret_cast<double2 *>(acc) + (size_t)mu * 4 * g.v + idx;
    q[0] = make_double2(z.a0, z.a1);
#pragma unroll
    for (int k = 0; k < 3; k++) q[(size_t)(k + 1) * g.v] = make_double2(z.p[2 * k], z.p[2 * k + 1]);
}
__device__ __forceinline__ su3alg<double> acc_load(const double *acc, const Geom &g, int mu, int idx)
{
    const double2 *q = reinterpret_cast<const double2 *>(acc) + (size_t)mu * 4 * g.v + idx;
    su3alg<double> z;
    const double2 u = q[0];
    z.a0 = u.x, z.a1 = u.y;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double2 v = q[(size_t)(k + 1) * g.v];
        z.p[2 * k] = v.x, z.p[2 * k + 1] = v.y;
    }
    return z;
}

template <class T, int R, int STAGE>
__global__ __launch_bounds__(kBlock) void flow_stage_kernel(const typename vec2<T>::type *__restrict__ win,
                                                            typename vec2<T>::type *__restrict__ wout,
                                                            T *__restrict__ acc, Geom g, T cz, T ca)
{
    __shared__ FlowLds<T> sh;
    const int mu = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // one direction per wavefront
    const int lane = threadIdx.x & 63;
    const int h_raw = xcd_block(blockIdx.x, gridDim.x) * (kFlowSites / 2) + (lane & 31);
    const bool valid = h_raw < g.vh;
    const int h = valid ? h_raw : g.vh - 1;  // idle lanes of a partial block mirror a real site, store nothing
    const int idx = parity_index(g, lane >> 5, h);
    const Coord x = decode(g, idx);
    const DirView<T> vmu = dir_view<T>(win, g, mu);
    lds_put(sh, mu, lane, load_link<T, R>(vmu, idx));
    lds_put(sh, 4 + mu, lane, load_link<T, R>(vmu, index_of(g, step(g, x, mu, -1))));
    __syncthreads();
    // staple A_mu(x) = sum_{nu != mu} [ U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
    //                                 + U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu) ]
    // (GLAC/actions/wilsongaugeaction.cpp:29-53)
    M3<T> a = zero3<T>();
    {
        const Coord xpm = step(g, x, mu, +1);
        const int ipm = index_of(g, xpm);
#pragma unroll 1
        for (int d = 1; d < 4; d++) {
            const int nu = (mu + d) & 3;
            const DirView<T> vnu = dir_view<T>(win, g, nu);
            const int ipn = index_of(g, step(g, x, nu, +1));
            const int imn = index_of(g, step(g, x, nu, -1));
            const int ipmmn = index_of(g, step(g, xpm, nu, -1));
            if constexpr (R == 2) {
                // Two stored rows: form each SU(3) product's first two rows
                // from operands of which only those are needed, and rebuild
                // the product's third row once (36 multiply-adds fewer per
                // product than rebuilding both operands).
                M3<T> p;  // U_nu(x+mu) U_mu(x+nu)^dag
                mm<false, true, false, 2>(p, load_rows01<T, R>(vnu, ipm), load_link<T, R>(vmu, ipn));
                rebuild_third_row(p);
                mm<false, true, true>(a, p, lds_get(sh, nu, lane));
                M3<T> q;  // U_mu(x-nu) U_nu(x+mu-nu), whose dagger the lower staple uses
                mm<false, false, false, 2>(q, load_rows01<T, R>(vmu, imn), load_link<T, R>(vnu, ipmmn));
                rebuild_third_row(q);
                mm<true, false, true>(a, q, lds_get(sh, 4 + nu, lane));
            } else {
                {
                    const M3<T> u1 = load_link<T, R>(vnu, ipm);
                    const M3<T> u2 = load_link<T, R>(vmu, ipn);
                    mm<false, true, true>(a, mul_ad(u1, u2), lds_get(sh, nu, lane));
                }
                {
                    const M3<T> u1 = load_link<T, R>(vnu, ipmmn);
                    const M3<T> u2 = load_link<T, R>(vmu, imn);
                    mm<false, false, true>(a, mul_dd(u1, u2), lds_get(sh, 4 + nu, lane));
                }
            }
        }
    }
    if (!valid) return;
    const M3<T> u = lds_get(sh, mu, lane);
    const su3alg<T> z = force_from_omega(mul(u, a));
    su3alg<T> y;
    if (STAGE == 0) {
        acc_store(acc, g, mu, idx, z);
        y.a0 = cz * z.a0;
        y.a1 = cz * z.a1;
#pragma unroll
        for (int k = 0; k < 6; k++) y.p[k] = cz * z.p[k];
    } else {
        const su3alg<T> c = acc_load(acc, g, mu, idx);
        y.a0 = cz * z.a0 - ca * c.a0;
        y.a1 = cz * z.a1 - ca * c.a1;
#pragma unroll
        for (int k = 0; k < 6; k++) y.p[k] = cz * z.p[k] - ca * c.p[k];
        if (STAGE == 1) acc_store(acc, g, mu, idx, y);
    }
    M3<T> w = zero3<T>();
    mm<false, false, false, R>(w, exp_su3(y), u);  // two rows suffice when two are stored
    if constexpr (sizeof(T) == 4) {
        // FP32 links: project back onto SU(3) in FP64 every stage.  Without
        // it the rounding of exp(Y) W drifts the links off the group
        // (|U U^dag - 1| ~ 1e-4 after 1000 steps at 24^3 x 48), and that drift,
        // not the FP32 measurement, dominates the error of E(t) at large t.
        M3<double> d = convert<double>(w);
        reunitarize<R>(d);
        w = convert<T>(d);
    }
    store_link<T, R>(wout, g, mu, idx, w);
}

// Dynamic LDS reserved per block; nonzero values cap the blocks resident per
// CU (a tuning parameter of the launch, set at build time).
#ifndef GLACX_FLOW_LDS_BYTES
#define GLACX_FLOW_LDS_BYTES 0
#endif

template <class T, int R>
void launch_flow_stage(const void *win, void *wout, void *acc, const Geom &g, int stage, double eps, hipStream_t s)
{
    using V2 = typename vec2<T>::type;
    const unsigned blocks = (unsigned)((g.vh + kFlowSites / 2 - 1) / (kFlowSites / 2));
    const V2 *in = (const V2 *)win;
    V2 *out = (V2 *)wout;
    T *a = (T *)acc;
    switch (stage) {
    case 0:
        flow_stage_kernel<T, R, 0><<<blocks, kBlock, GLACX_FLOW_LDS_BYTES, s>>>(in, out, a, g, T(0.25 * eps), T(0));
        break;
    case 1:
        flow_stage_kernel<T, R, 1><<<blocks, kBlock, GLACX_FLOW_LDS_BYTES, s>>>(in, out, a, g, T(8.0 / 9.0 * eps), T(17.0 / 36.0 * eps));
        break;
    default:
        flow_stage_kernel<T, R, 2><<<blocks, kBlock, GLACX_FLOW_LDS_BYTES, s>>>(in, out, a, g, T(0.75 * eps), T(1));
        break;
    }
}

template void launch_flow_stage<float, 3>(const void *, void *, void *, const Geom &, int, double, hipStream_t);
template void launch_flow_stage<float, 2>(const void *, void *, void *, const Geom &, int, double, hipStream_t);
template void launch_flow_stage<double, 3>(const void *, void *, void *, const Geom &, int, double, hipStream_t);

} // namespace glacx
