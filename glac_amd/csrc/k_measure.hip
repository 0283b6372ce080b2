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

// op(A) op(B) op(C) op(D) of four SU(3) factors (op = dagger where the flag
// is set): each partial product keeps only its first two rows, which need
// only the first two rows of the partial product before it, and the third
// row of the result is rebuilt at the end -- 3 x 72 multiply-adds and one
// rebuild instead of 3 x 108.
// The rebuild holds for det = +1 only, so it is used with two stored rows
// (R = 2, whose links are checked to be SU(3)); three stored rows may carry
// GLAC's det = -1 links and take the full products.
template <bool DA, bool DB, bool DC, bool DD, int R, class T>
__device__ __forceinline__ M3<T> su3_chain4(const M3<T> &a, const M3<T> &b, const M3<T> &c, const M3<T> &d)
{
    M3<T> p, q;
    if constexpr (R == 2) {
        mm<DA, DB, false, 2>(p, a, b);
        mm<false, DC, false, 2>(q, p, c);
        mm<false, DD, false, 2>(p, q, d);
        rebuild_third_row(p);
    } else {
        mm<DA, DB, false>(p, a, b);
        mm<false, DC, false>(q, p, c);
        mm<false, DD, false>(p, q, d);
    }
    return p;
}

// F = T(C) of the clover C = L1 - L2 + L3 - L4 of plane (mu, nu) at x, in
// compact form: T is linear, so F = T(L1) - T(L2) + T(L3) - T(L4), and each
// leaf's last product forms only the entries T needs.  Accumulating 8 reals
// instead of a 3 x 3 complex C keeps fewer registers live.  plaq = Re tr L1.
template <class T, int R>
__device__ __forceinline__ void clover(const typename vec2<T>::type *__restrict__ f, const Geom &g, const Coord &x,
                                       int mu, int nu, T &plaq, T F[8])
{
    const Coord xpm = step(g, x, mu, +1), xpn = step(g, x, nu, +1);
    const Coord xmm = step(g, x, mu, -1), xmn = step(g, x, nu, -1);
    const Coord xpmmn = step(g, xpm, nu, -1), xmmmn = step(g, xmm, nu, -1), xmmpn = step(g, xmm, nu, +1);
    const int ix = index_of(g, x), imn = index_of(g, xmn), imm = index_of(g, xmm);
    const DirView<T> vmu = dir_view<T>(f, g, mu), vnu = dir_view<T>(f, g, nu);
    const M3<T> unu = load_link<T, R>(vnu, ix);
    const int immmn = index_of(g, xmmmn);
    T t[8];
    if constexpr (R == 2) {
        const M3<T> unu_mn = load_link<T, R>(vnu, imn);
        const M3<T> umu_mm = load_link<T, R>(vmu, imm);
        const M3<T> umu = load_rows01<T, R>(vmu, ix);  // a leading factor only: two rows
        // L1 = U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
        {
            const M3<T> c = su3_chain4<false, false, true, true, R>(umu, load_link<T, R>(vnu, index_of(g, xpm)),
                                                                    load_link<T, R>(vmu, index_of(g, xpn)), unu);
            plaq = retrace(c);
            ta2_of(c, F);
        }
        // - L2 = -U_mu(x) U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu)
        ta2_of(su3_chain4<false, true, true, false, R>(umu, load_link<T, R>(vnu, index_of(g, xpmmn)),
                                                       load_link<T, R>(vmu, imn), unu_mn),
               t);
#pragma unroll
        for (int k = 0; k < 8; k++) F[k] -= t[k];
        // + L3 = U_mu(x-mu)^dag U_nu(x-mu-nu)^dag U_mu(x-mu-nu) U_nu(x-nu)
        ta2_of(su3_chain4<true, true, false, false, R>(umu_mm, load_link<T, R>(vnu, immmn),
                                                       load_link<T, R>(vmu, immmn), unu_mn),
               t);
#pragma unroll
        for (int k = 0; k < 8; k++) F[k] += t[k];
        // - L4 = -U_mu(x-mu)^dag U_nu(x-mu) U_mu(x-mu+nu) U_nu(x)^dag
        ta2_of(su3_chain4<true, false, false, true, R>(umu_mm, load_link<T, R>(vnu, imm),
                                                       load_link<T, R>(vmu, index_of(g, xmmpn)), unu),
               t);
#pragma unroll
        for (int k = 0; k < 8; k++) F[k] -= t[k];
    } else {
        // full products (three stored rows may hold GLAC's det = -1 links)
        const M3<T> umu = load_link<T, R>(vmu, ix);
        {
            const M3<T> c =
                mul_ad(mul_ad(mul(umu, load_link<T, R>(vnu, index_of(g, xpm))), load_link<T, R>(vmu, index_of(g, xpn))),
                       unu);
            plaq = retrace(c);
            ta2_of(c, F);
        }
        const M3<T> unu_mn = load_link<T, R>(vnu, imn);
        ta2_product<false, false>(mul_ad(mul_ad(umu, load_link<T, R>(vnu, index_of(g, xpmmn))), load_link<T, R>(vmu, imn)),
                                  unu_mn, t);
#pragma unroll
        for (int k = 0; k < 8; k++) F[k] -= t[k];
        const M3<T> umu_mm = load_link<T, R>(vmu, imm);
        ta2_product<false, false>(mul(mul_dd(umu_mm, load_link<T, R>(vnu, immmn)), load_link<T, R>(vmu, immmn)), unu_mn,
                                  t);
#pragma unroll
        for (int k = 0; k < 8; k++) F[k] += t[k];
        ta2_product<false, true>(mul(mul_da(umu_mm, load_link<T, R>(vnu, imm)), load_link<T, R>(vmu, index_of(g, xmmpn))),
                                 unu, t);
#pragma unroll
        for (int k = 0; k < 8; k++) F[k] -= t[k];
    }
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

// LDS exchange of the six field-strength tensors of a block's 64 sites
// (SuperSampler variant): slot-major, then matrix entry, then site, so the
// 64 lanes of a wavefront touch consecutive words.
template <class T> struct MeasLds {
    T v[6][18][kMeasSites];
};
template <class T> __device__ __forceinline__ MeasLds<T> &meas_lds()
{
    __shared__ MeasLds<T> s;
    return s;
}
template <class T> __device__ __forceinline__ void lds_put(int slot, int lane, const M3<T> &m)
{
    MeasLds<T> &s = meas_lds<T>();
#pragma unroll
    for (int e = 0; e < 9; e++) {
        s.v[slot][2 * e][lane] = m.e[e / 3][e % 3].r;
        s.v[slot][2 * e + 1][lane] = m.e[e / 3][e % 3].i;
    }
}
template <class T> __device__ __forceinline__ M3<T> lds_get(int slot, int lane)
{
    const MeasLds<T> &s = meas_lds<T>();
    M3<T> m;
#pragma unroll
    for (int e = 0; e < 9; e++) m.e[e / 3][e % 3] = {s.v[slot][2 * e][lane], s.v[slot][2 * e + 1][lane]};
    return m;
}

// SUPER adds GLAC's SuperSampler Weinberg operator
// (GLAC/observables/supersampler.cpp:430-460): per site
//   w = -sum_{nu=1..3} sum_{lambda != 0, nu} Re tr(G_{0 lambda} G_{nu lambda} G_{rho sigma}).
// Wavefront nu holds G_{0 nu} and G_{rho sigma} for the block's sites; the
// six tensors are exchanged through LDS (slot nu - 1: G_{0 nu}, slot 2 + nu:
// G_{rho sigma}, i.e. G_23, G_31, G_12 in slots 3, 4, 5), then wavefront nu
// forms its two terms with lambda_a = next(nu), lambda_b = next(lambda_a):
// G_{nu lambda_a} is stored as such (slot 2 + lambda_b) and
// G_{nu lambda_b} = G_{lambda_b nu}^dagger (slot 2 + lambda_a).
//
// DENSITY also writes the per-site charge and action densities of GLAC's
// LatticeActionChargeDensity (GLAC/observables/latticeactionchargedensity.cpp:
// 169-306): q(x) = -sum_nu Re tr(G_{0 nu} G_{rho sigma}) / (256 pi^2) and
// e(x) = sum_nu [Re tr(G_{0 nu}^2) + Re tr(G_{rho sigma}^2)], at GLAC's
// lexicographic site index; the three wavefronts' terms of a site are
// added through LDS in the order nu = 1, 2, 3.
enum MeasMode { MEAS_PLAIN = 0, MEAS_SUPER = 1, MEAS_DENSITY = 2 };

// Waves per SIMD the clover kernel is compiled for (build-time tuning): FP64
// needs ~250 VGPRs (two waves); FP32 takes ~145 at three waves, and four
// (128 VGPRs, a few spilled) measured no faster.
#ifndef GLACX_MEAS_FP32_WAVES
#define GLACX_MEAS_FP32_WAVES 3
#endif
template <class T> constexpr int meas_waves() { return sizeof(T) == 8 ? 2 : GLACX_MEAS_FP32_WAVES; }

template <class T, int R, int MODE>
__global__ __launch_bounds__(kMeasSites * kMeasWaves)
__attribute__((amdgpu_waves_per_eu(meas_waves<T>(), meas_waves<T>()))) void measure_kernel(const typename vec2<T>::type *__restrict__ f,
                                                                         Geom g, int cps, double *__restrict__ partials,
                                                                         double *__restrict__ density)
{
    constexpr bool SUPER = MODE == MEAS_SUPER, DENSITY = MODE == MEAS_DENSITY;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int rest = fdiv(b, g.d_cpr);
    int jc = b - rest * g.cpr, lt = rest & ((1 << g.lbt) - 1);
    const int blk = rest >> g.lbt;
    const int plane = g.run >> g.lbz;  // Ny Nx / 2 entries per x-y plane of one parity
    if ((plane & 31) == 0) {
        // visit the slab's chunks y-chunk-major, t fastest (as the flow stage
        // does): z- and t-neighbours then at most b_z b_t chunks apart
        const int r = b - (blk << g.lbt) * g.cpr, q = r >> g.lbt;
        lt = r & ((1 << g.lbt) - 1);
        jc = (q & ((1 << g.lbz) - 1)) * (plane >> 5) + (q >> g.lbz);
    }
    const int bt = fdiv(blk, g.d_nzb), zb = blk - bt * g.nzb;
    const int t = (bt << g.lbt) + lt, c = zb * g.cpr + jc;
    const int lane = threadIdx.x & 63;
    const int k = jc * (kMeasSites / 2) + (lane & 31);
    const int nu = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) + 1;  // one clover pair per wavefront
    const bool valid = k < g.run;
    double acc[kMeasAcc] = {0.0, 0.0, 0.0, 0.0};
    M3<T> c2;
    int lex = 0;
    if (valid) {
        const Coord x = decode(g, run_index(g, blk, lane >> 5, lt, k));
        lex = lex_of(g, x);
        const int rho = nu % 3 + 1, sigma = rho % 3 + 1;
        T p1, p2, f1[8], f2[8];
        clover<T, R>(f, g, x, 0, nu, p1, f1);
        clover<T, R>(f, g, x, rho, sigma, p2, f2);
        acc[0] = p1 + p2;
        acc[1] = retrace_ta2(f1, f1) + retrace_ta2(f2, f2);
        acc[2] = -retrace_ta2(f1, f2);
        if constexpr (SUPER) {
            c2 = ta2_matrix(f2);
            lds_put(nu - 1, lane, ta2_matrix(f1));
            lds_put(2 + nu, lane, c2);
        }
    }
    if constexpr (SUPER) {
        __syncthreads();
        if (valid) {
            const int la = nu % 3 + 1, lb = la % 3 + 1;
            const double ta = retrace_mul(mul(lds_get<T>(la - 1, lane), lds_get<T>(2 + lb, lane)), c2);
            const double tb = retrace_mul(mul_ad(lds_get<T>(lb - 1, lane), lds_get<T>(2 + la, lane)), c2);
            acc[3] = -(ta + tb);
        }
    }
    if constexpr (DENSITY) {
        __shared__ double dens[2][kMeasWaves][kMeasSites];
        dens[0][nu - 1][lane] = acc[2];
        dens[1][nu - 1][lane] = acc[1];
        __syncthreads();
        if (nu == 1 && valid) {
            const double qnorm = 1.0 / (256.0 * M_PI * M_PI);
            density[lex] = ((dens[0][0][lane] + dens[0][1][lane]) + dens[0][2][lane]) * qnorm;
            density[(size_t)g.v + lex] = (dens[1][0][lane] + dens[1][1][lane]) + dens[1][2][lane];
        }
    }
    block_sum<kMeasAcc, kMeasWaves>(acc);
    if (threadIdx.x == 0) {
        double *q = partials + kMeasAcc * (c * g.n[3] + t);
#pragma unroll
        for (int i = 0; i < kMeasAcc; i++) q[i] = acc[i];
    }
}

// One block per time slice t: sums the partials of the chunks of slice t in
// a fixed order into slice[kMeasAcc t ...].
__global__ __launch_bounds__(kBlock) void measure_slice_kernel(const double *__restrict__ partials, Geom g, int cps,
                                                               double *__restrict__ slice)
{
    const int t = blockIdx.x;
    double acc[kMeasAcc] = {0.0, 0.0, 0.0, 0.0};
    for (int c = threadIdx.x; c < cps; c += kBlock) {
        const double *q = partials + kMeasAcc * (c * g.n[3] + t);
#pragma unroll
        for (int i = 0; i < kMeasAcc; i++) acc[i] += q[i];
    }
    block_sum<kMeasAcc>(acc);
    if (threadIdx.x == 0)
#pragma unroll
        for (int i = 0; i < kMeasAcc; i++) slice[kMeasAcc * t + i] = acc[i];
}

// out = {plaq, E_raw, Q, Q(t = 0..Nt-1)}, or with `super`
// {plaq, E_raw, Q, W, Q(t)..., W(t)...}; W's normalisation is GLAC's
// 2 * 8 * 8 / (3 * 16^3) (GLAC/observables/supersampler.cpp:17-19).
// One wavefront: lane l takes the slices t = l, l + 64, ... in order, then
// the lanes' sums are combined by a fixed shuffle tree (the order never
// depends on scheduling).
__global__ __launch_bounds__(64) void measure_finish_kernel(const double *__restrict__ slice, Geom g, int super,
                                                            double *__restrict__ out)
{
    const double qnorm = 1.0 / (256.0 * M_PI * M_PI), wnorm = 2.0 * 8.0 * 8.0 / (3.0 * 16.0 * 16.0 * 16.0);
    const int nt = g.n[3], head = super ? 4 : 3, lane = threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};  // plaquette, energy, q, w
    for (int t = lane; t < nt; t += 64) {
        const double *st = slice + kMeasAcc * t;
        v[0] += st[0];
        v[1] += st[1];
        const double qt = st[2] * qnorm;
        out[head + t] = qt;
        v[2] += qt;
        if (super) {
            const double wt = st[3] * wnorm;
            out[head + nt + t] = wt;
            v[3] += wt;
        }
    }
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
    if (lane == 0) {
        out[0] = v[0] / (18.0 * (double)g.v);
        out[1] = v[1] / (double)g.v;
        out[2] = v[2];
        if (super) out[3] = v[3];
    }
}

// Plaquette only: sum over sites and the 6 planes of
// Re tr[U_mu(x) U_nu(x+mu) (U_nu(x) U_mu(x+nu))^dagger]
// (GLAC/observables/plaquette.cpp:26-40).
template <class T, int R>
__global__ __launch_bounds__(kBlock) void plaquette_kernel(const typename vec2<T>::type *__restrict__ f, Geom g,
                                                           double *__restrict__ partials)
{
    const int idx = blockIdx.x * kBlock + threadIdx.x;
    double acc[1] = {0.0};
    if (idx < g.v) {
        const Site x = site_of(g, idx);
        T s = T(0);
#pragma unroll 1
        for (int mu = 0; mu < 3; mu++) {
            const M3<T> umu = load_link<T, R>(f, g, mu, idx);
            const int ipm = site_step(g, x, mu, +1).idx;
#pragma unroll 1
            for (int nu = mu + 1; nu < 4; nu++) {
                const M3<T> a = mul(umu, load_link<T, R>(f, g, nu, ipm));
                const M3<T> b = mul(load_link<T, R>(f, g, nu, idx), load_link<T, R>(f, g, mu, site_step(g, x, nu, +1).idx));
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
int measure_partials_count(const Geom &g) { return kMeasAcc * g.n[3] * chunks_per_slab(g) + kMeasAcc * g.n[3]; }
int plaquette_partials_count(const Geom &g) { return (g.v + kBlock - 1) / kBlock; }

template <class T, int R>
void launch_measure(const void *f, const Geom &g, double *partials, int mode, double *density, hipStream_t s)
{
    const int cps = chunks_per_slab(g);
    const auto *fp = (const typename vec2<T>::type *)f;
    const unsigned blocks = (unsigned)(g.n[3] * cps), threads = kMeasSites * kMeasWaves;
    if (mode == MEAS_SUPER)
        measure_kernel<T, R, MEAS_SUPER><<<blocks, threads, 0, s>>>(fp, g, cps, partials, density);
    else if (mode == MEAS_DENSITY)
        measure_kernel<T, R, MEAS_DENSITY><<<blocks, threads, 0, s>>>(fp, g, cps, partials, density);
    else
        measure_kernel<T, R, MEAS_PLAIN><<<blocks, threads, 0, s>>>(fp, g, cps, partials, density);
}

void launch_measure_finish(double *partials, const Geom &g, double *out, bool super, hipStream_t s)
{
    const int cps = chunks_per_slab(g);
    double *slice = partials + kMeasAcc * g.n[3] * cps;
    measure_slice_kernel<<<g.n[3], kBlock, 0, s>>>(partials, g, cps, slice);
    measure_finish_kernel<<<1, 64, 0, s>>>(slice, g, super ? 1 : 0, out);
}

template <class T, int R> void launch_plaquette(const void *f, const Geom &g, double *partials, hipStream_t s)
{
    plaquette_kernel<T, R><<<plaquette_partials_count(g), kBlock, 0, s>>>((const typename vec2<T>::type *)f, g, partials);
}

void launch_plaquette_finish(const double *partials, int nparts, const Geom &g, double *out, hipStream_t s)
{
    plaquette_finish_kernel<<<1, kBlock, 0, s>>>(partials, nparts, g, out);
}

template void launch_measure<float, 3>(const void *, const Geom &, double *, int, double *, hipStream_t);
template void launch_measure<float, 2>(const void *, const Geom &, double *, int, double *, hipStream_t);
template void launch_measure<double, 3>(const void *, const Geom &, double *, int, double *, hipStream_t);
template void launch_plaquette<float, 3>(const void *, const Geom &, double *, hipStream_t);
template void launch_plaquette<float, 2>(const void *, const Geom &, double *, hipStream_t);
template void launch_plaquette<double, 3>(const void *, const Geom &, double *, hipStream_t);

} // namespace glacx
