// k_metropolis.hip -- one checkerboard Metropolis sub-sweep (mu, parity).
//
// Replaces System::update / updateLink (GLAC/system.cpp:321-361): GLAC walks
// the links sequentially; here every link U_mu(x) with x of the given parity
// is updated at once.  Its staple only involves U_mu at the other parity and
// U_nu, nu != mu, none of which change during the sub-sweep, so the links are
// independent and are updated in place.
//
// Per link: the staple A (WilsonGaugeAction::computeStaple,
// GLAC/actions/wilsongaugeaction.cpp:29-53), then n_hit proposals U' = X U
// with X from GLAC's RST construction (GLAC/math/matrices/
// su3matrixgenerator.h:170-321): three SU(2) matrices x0 1 + i x.sigma in
// blocks (0,1), (0,2), (1,2), x0 = sign(r0) sqrt(1 - eps^2),
// x = eps r / |r|, r ~ U(-1/2, 1/2)^4, product R S T or its inverse by a coin.
// dS = -(beta/3) Re tr[(U' - U) A] (wilsongaugeaction.cpp:9, 24-27) is
// evaluated as -(beta/3) (Re tr(U' A) - Re tr(U A)), with Re tr(U A) carried
// from the last accepted hit; accept iff exp(-dS) > u (GLAC/system.cpp:347).
// After the hits the link is optionally reunitarised in FP64, keeping the
// sign of det U (three stored rows may hold GLAC's det = -1 links).  Uniforms
// come from the link's own counter-based stream (rng.hpp): per hit, GLAC's 13
// centred ones for the proposal as 16-bit halves of seven outputs, then one
// 32-bit open one for the test.
#include "kernels.hpp"
#include "proposal.hpp"

namespace glacx {

template <class T, int R>
__global__ __launch_bounds__(kBlock, sizeof(T) == 8 ? 2 : 1) void metropolis_kernel(typename vec2<T>::type *__restrict__ f, Geom g, int mu,
                                                            int parity, MetroParams prm,
                                                            unsigned long long *__restrict__ count)
{
    const int h = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    double nacc[1] = {0.0};
    if (h < g.vh) {
        const int idx = parity_index(g, parity, h);
        const Site x = site_of(g, idx);
        const M3<T> a = staple<T, R>(f, g, x, mu);
        M3<T> u = load_link<T, R>(f, g, mu, idx);
        Stream rs(prm.seed, (uint32_t)lex_of(g, x), (uint32_t)mu, STREAM_METROPOLIS, prm.sweep);
        const T eps = T(prm.eps);
        const T x0abs = sqrt(T(1) - eps * eps);
        const double bf = prm.beta / 3.0;
        T tra = retrace_mul(u, a);
        int accepted = 0;
#pragma unroll 1
        for (int hit = 0; hit < prm.n_hit; hit++) {
            const M3<T> v = propose_link<T, R>(rst_proposal<T>(rs, x0abs, eps), u);
            const T trv = retrace_mul(v, a);
            const T ds = T(-bf) * (trv - tra);
            if (accept_prob(ds) > rs.open()) {
                u = v;
                tra = trv;
                accepted++;
            }
        }
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
    block_sum<1>(nacc);
    if (threadIdx.x == 0) atomicAdd(count, (unsigned long long)nacc[0]);
}

template <class T, int R>
void launch_metropolis(void *f, const Geom &g, int mu, int parity, const MetroParams &p, unsigned long long *count,
                       hipStream_t s)
{
    const unsigned blocks = (unsigned)((g.vh + kBlock - 1) / kBlock);
    metropolis_kernel<T, R><<<blocks, kBlock, 0, s>>>((typename vec2<T>::type *)f, g, mu, parity, p, count);
}

template void launch_metropolis<float, 3>(void *, const Geom &, int, int, const MetroParams &, unsigned long long *,
                                       hipStream_t);
template void launch_metropolis<float, 2>(void *, const Geom &, int, int, const MetroParams &, unsigned long long *,
                                       hipStream_t);
template void launch_metropolis<double, 3>(void *, const Geom &, int, int, const MetroParams &, unsigned long long *,
                                        hipStream_t);

} // namespace glacx
