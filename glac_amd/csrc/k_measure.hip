// k_measure.hip -- clover observables (plaquette, E_raw, Q, Q(t)) and the
// plaquette-only sum.
//
// Restates MasterSamplerTopcXYZ::calculate
// (GLAC/observables/mastersamplertopcxyz.cpp:236-384): for (mu, nu) =
// (0, nu), nu = 1..3, paired with (rho, sigma) = (nu % 3 + 1, rho % 3 + 1),
// the four-leaf clover C = L1 - L2 + L3 - L4 with GLAC's leaf orientations
// (:255-328), C <- traceless(C - C^dagger) (:331-340), then per site
//   plaq += Re tr L1 (both planes),  E += Re tr(C1 C1) + Re tr(C2 C2),
//   q -= Re tr(C1 C2).
// Normalisations (:14-16): plaq / (18 V), E / V, Q / (256 pi^2).
//
// Each block covers 64 sites of one time slice, so its q partial belongs to
// one slice; per-site values are summed in FP64, the block partials are
// combined per time slice and then over slices in a fixed order, which makes
// the result independent of scheduling.
#include "kernels.hpp"

namespace glacx {

template <class T>
__device__ __forceinline__ M3<T> clover(const typename vec2<T>::type *__restrict__ f, const Geom &g, const Coord &x,
                                        int mu, int nu, T &plaq)
{
    const Coord xpm = step(g, x, mu, +1), xpn = step(g, x, nu, +1);
    const Coord xmm = step(g, x, mu, -1), xmn = step(g, x, nu, -1);
    const Coord xpmmn = step(g, xpm, nu, -1), xmmmn = step(g, xmm, nu, -1), xmmpn = step(g, xmm, nu, +1);
    const int ix = index_of(g, x), imn = index_of(g, xmn), imm = index_of(g, xmm);
    const DirView<T> vmu = dir_view<T>(f, g.v, mu), vnu = dir_view<T>(f, g.v, nu);
    const M3<T> umu = load_link<T>(vmu, ix);
    const M3<T> unu = load_link<T>(vnu, ix);
    // L1 = U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
    M3<T> c = mul_ad(mul_ad(mul(umu, load_link<T>(vnu, index_of(g, xpm))), load_link<T>(vmu, index_of(g, xpn))), unu);
    plaq = retrace(c);
    // - L2 = -U_mu(x) U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu)
    const M3<T> unu_mn = load_link<T>(vnu, imn);
    {
        M3<T> m = mul_ad(mul_ad(umu, load_link<T>(vnu, index_of(g, xpmmn))), load_link<T>(vmu, imn));
        negate(m);
        mm<false, false, true>(c, m, unu_mn);
    }
    // + L3 = U_mu(x-mu)^dag U_nu(x-mu-nu)^dag U_mu(x-mu-nu) U_nu(x-nu)
    const M3<T> umu_mm = load_link<T>(vmu, imm);
    const int immmn = index_of(g, xmmmn);
    mm<false, false, true>(c, mul(mul_dd(umu_mm, load_link<T>(vnu, immmn)), load_link<T>(vmu, immmn)), unu_mn);
    // - L4 = -U_mu(x-mu)^dag U_nu(x-mu) U_mu(x-mu+nu) U_nu(x)^dag
    {
        M3<T> m = mul(mul_da(umu_mm, load_link<T>(vnu, imm)), load_link<T>(vmu, index_of(g, xmmpn)));
        negate(m);
        mm<false, true, true>(c, m, unu);
    }
    return c;
}

// Launch shape: 192-thread blocks = 64 sites x 3 wavefronts; wavefront w
// evaluates the clover pair nu = w + 1 (planes (0, nu) and (rho, sigma)) for
// the block's 64 sites, so the three pairs of a site run side by side.
// Blocks follow the storage order: block b takes chunk jc = b % cpr of the
// run (blk, lt), i.e. its sites jc * 32 .. jc * 32 + 31 of both parities
// (neighbours of each other), all on time slice t = (blk / nzb) bt + lt.
// Its partials go to slot (c, t), c = zb cpr + jc being the chunk's position
// within its time slice, where the per-slice sum finds them in a fixed order.
constexpr int kMeasSites = 64;
constexpr int kMeasWaves = 3;

template <class T>
__global__ __launch_bounds__(kMeasSites *kMeasWaves) void measure_kernel(const typename vec2<T>::type *__restrict__ f,
                                                                         Geom g, int cps, double *__restrict__ partials)
{
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int rest = fdiv(b, g.d_cpr), jc = b - rest * g.cpr;
    const int lt = rest & ((1 << g.lbt) - 1), blk = rest >> g.lbt;
    const int bt = fdiv(blk, g.d_nzb), zb = blk - bt * g.nzb;
    const int t = (bt << g.lbt) + lt, c = zb * g.cpr + jc;
    const int lane = threadIdx.x & 63;
    const int k = jc * (kMeasSites / 2) + (lane & 31);
    const int nu = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) + 1;  // one clover pair per wavefront
    double acc[3] = {0.0, 0.0, 0.0};
    if (k < g.run) {
        const Coord x = decode(g, run_index(g, blk, lane >> 5, lt, k));
        const int rho = nu % 3 + 1, sigma = rho % 3 + 1;
        T p1, p2;
        const M3<T> c1 = traceless_antiherm2(clover<T>(f, g, x, 0, nu, p1));
        const M3<T> c2 = traceless_antiherm2(clover<T>(f, g, x, rho, sigma, p2));
        acc[0] = p1 + p2;
        acc[1] = retrace_mul(c1, c1) + retrace_mul(c2, c2);
        acc[2] = -retrace_mul(c1, c2);
    }
    block_sum<3, kMeasWaves>(acc);
    if (threadIdx.x == 0) {
        double *q = partials + 3 * (c * g.n[3] + t);
        q[0] = acc[0];
        q[1] = acc[1];
        q[2] = acc[2];
    }
}

// One block per time slice t: sums the partials of the chunks of slice t in
// a fixed order into slice[3 t .. 3 t + 2].
__global__ __launch_bounds__(kBlock) void measure_slice_kernel(const double *__restrict__ partials, Geom g, int cps,
                                                               double *__restrict__ slice)
{
    const int t = blockIdx.x;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int c = threadIdx.x; c < cps; c += kBlock) {
        const double *q = partials + 3 * (c * g.n[3] + t);
        acc[0] += q[0];
        acc[1] += q[1];
        acc[2] += q[2];
    }
    block_sum<3>(acc);
    if (threadIdx.x == 0) {
        slice[3 * t + 0] = acc[0];
        slice[3 * t + 1] = acc[1];
        slice[3 * t + 2] = acc[2];
    }
}

// out = {plaq, E_raw, Q, Q(t = 0..Nt-1)}.
__global__ void measure_finish_kernel(const double *__restrict__ slice, Geom g, double *__restrict__ out)
{
    if (threadIdx.x != 0) return;
    const double qnorm = 1.0 / (256.0 * M_PI * M_PI);
    double pl = 0.0, en = 0.0, q = 0.0;
    for (int t = 0; t < g.n[3]; t++) {
        pl += slice[3 * t + 0];
        en += slice[3 * t + 1];
        const double qt = slice[3 * t + 2] * qnorm;
        out[3 + t] = qt;
        q += qt;
    }
    out[0] = pl / (18.0 * (double)g.v);
    out[1] = en / (double)g.v;
    out[2] = q;
}

// Plaquette only: sum over sites and the 6 planes of
// Re tr[U_mu(x) U_nu(x+mu) (U_nu(x) U_mu(x+nu))^dagger]
// (GLAC/observables/plaquette.cpp:26-40).
template <class T>
__global__ __launch_bounds__(kBlock) void plaquette_kernel(const typename vec2<T>::type *__restrict__ f, Geom g,
                                                           double *__restrict__ partials)
{
    const int idx = blockIdx.x * kBlock + threadIdx.x;
    double acc[1] = {0.0};
    if (idx < g.v) {
        const Coord x = decode(g, idx);
        T s = T(0);
#pragma unroll 1
        for (int mu = 0; mu < 3; mu++) {
            const M3<T> umu = load_link<T>(f, g.v, mu, idx);
            const int ipm = index_of(g, step(g, x, mu, +1));
#pragma unroll 1
            for (int nu = mu + 1; nu < 4; nu++) {
                const M3<T> a = mul(umu, load_link<T>(f, g.v, nu, ipm));
                const M3<T> b = mul(load_link<T>(f, g.v, nu, idx), load_link<T>(f, g.v, mu, index_of(g, step(g, x, nu, +1))));
                // Re tr(a b^dag) = sum_ij Re(a_ij conj(b_ij))
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++) s += a.e[i][j].r * b.e[i][j].r + a.e[i][j].i * b.e[i][j].i;
            }
        }
        acc[0] = s;
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc[0];
}

__global__ __launch_bounds__(kBlock) void plaquette_finish_kernel(const double *__restrict__ partials, int n, Geom g,
                                                                  double *__restrict__ out)
{
    double acc[1] = {0.0};
    for (int b = threadIdx.x; b < n; b += kBlock) acc[0] += partials[b];
    block_sum<1>(acc);
    if (threadIdx.x == 0) out[0] = acc[0] / (18.0 * (double)g.v);
}

// 64-site chunks per time slice: nzb runs of each slice, cpr chunks per run
static inline int chunks_per_slab(const Geom &g) { return g.nzb * g.cpr; }

// block partials followed by the per-slice sums
int measure_partials_count(const Geom &g) { return 3 * g.n[3] * chunks_per_slab(g) + 3 * g.n[3]; }
int plaquette_partials_count(const Geom &g) { return (g.v + kBlock - 1) / kBlock; }

template <class T> void launch_measure(const void *f, const Geom &g, double *partials, hipStream_t s)
{
    const int cps = chunks_per_slab(g);
    measure_kernel<T><<<g.n[3] * cps, kMeasSites * kMeasWaves, 0, s>>>((const typename vec2<T>::type *)f, g, cps,
                                                                             partials);
}

void launch_measure_finish(double *partials, const Geom &g, double *out, hipStream_t s)
{
    const int cps = chunks_per_slab(g);
    double *slice = partials + 3 * g.n[3] * cps;
    measure_slice_kernel<<<g.n[3], kBlock, 0, s>>>(partials, g, cps, slice);
    measure_finish_kernel<<<1, 64, 0, s>>>(slice, g, out);
}

template <class T> void launch_plaquette(const void *f, const Geom &g, double *partials, hipStream_t s)
{
    plaquette_kernel<T><<<plaquette_partials_count(g), kBlock, 0, s>>>((const typename vec2<T>::type *)f, g, partials);
}

void launch_plaquette_finish(const double *partials, int nparts, const Geom &g, double *out, hipStream_t s)
{
    plaquette_finish_kernel<<<1, kBlock, 0, s>>>(partials, nparts, g, out);
}

template void launch_measure<float>(const void *, const Geom &, double *, hipStream_t);
template void launch_measure<double>(const void *, const Geom &, double *, hipStream_t);
template void launch_plaquette<float>(const void *, const Geom &, double *, hipStream_t);
template void launch_plaquette<double>(const void *, const Geom &, double *, hipStream_t);

} // namespace glacx
