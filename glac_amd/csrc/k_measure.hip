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
// Each block covers one (parity, time slice) slab chunk, so its q partial
// belongs to one time slice; per-site values are summed in FP64 and the
// partials are combined in a fixed order by a single-block finishing kernel,
// which makes the result independent of scheduling.
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
    const M3<T> umu = load_link<T>(f, g.v, mu, ix);
    const M3<T> unu = load_link<T>(f, g.v, nu, ix);
    // L1 = U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
    M3<T> c = mul_ad(mul_ad(mul(umu, load_link<T>(f, g.v, nu, index_of(g, xpm))),
                            load_link<T>(f, g.v, mu, index_of(g, xpn))),
                     unu);
    plaq = retrace(c);
    // L2 = U_mu(x) U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu)
    const M3<T> unu_mn = load_link<T>(f, g.v, nu, imn);
    sub_from(c, mul(mul_ad(mul_ad(umu, load_link<T>(f, g.v, nu, index_of(g, xpmmn))),
                           load_link<T>(f, g.v, mu, imn)),
                    unu_mn));
    // L3 = U_mu(x-mu)^dag U_nu(x-mu-nu)^dag U_mu(x-mu-nu) U_nu(x-nu)
    const M3<T> umu_mm = load_link<T>(f, g.v, mu, imm);
    const int immmn = index_of(g, xmmmn);
    add_to(c, mul(mul(mul_dd(umu_mm, load_link<T>(f, g.v, nu, immmn)), load_link<T>(f, g.v, mu, immmn)), unu_mn));
    // L4 = U_mu(x-mu)^dag U_nu(x-mu) U_mu(x-mu+nu) U_nu(x)^dag
    sub_from(c, mul_ad(mul(mul_da(umu_mm, load_link<T>(f, g.v, nu, imm)),
                           load_link<T>(f, g.v, mu, index_of(g, xmmpn))),
                       unu));
    return c;
}

template <class T>
__global__ __launch_bounds__(kBlock) void measure_kernel(const typename vec2<T>::type *__restrict__ f, Geom g, int cps,
                                                         double *__restrict__ partials)
{
    const int b = blockIdx.x;
    const int slab = b / cps, chunk = b - slab * cps;  // slab = parity * Nt + t
    const int p = slab / g.n[3], t = slab - p * g.n[3];
    const int r = chunk * kBlock + threadIdx.x;
    double acc[3] = {0.0, 0.0, 0.0};
    if (r < g.vs_h) {
        const Coord x = decode(g, p * g.vh + t * g.vs_h + r);
        T plaq = T(0), en = T(0), q = T(0);
#pragma unroll 1
        for (int nu = 1; nu < 4; nu++) {
            const int rho = nu % 3 + 1, sigma = rho % 3 + 1;
            T p1, p2;
            const M3<T> c1 = traceless_antiherm2(clover<T>(f, g, x, 0, nu, p1));
            const M3<T> c2 = traceless_antiherm2(clover<T>(f, g, x, rho, sigma, p2));
            plaq += p1 + p2;
            q -= retrace_mul(c1, c2);
            en += retrace_mul(c1, c1) + retrace_mul(c2, c2);
        }
        acc[0] = plaq;
        acc[1] = en;
        acc[2] = q;
    }
    block_sum<3>(acc);
    if (threadIdx.x == 0) {
        partials[3 * b + 0] = acc[0];
        partials[3 * b + 1] = acc[1];
        partials[3 * b + 2] = acc[2];
    }
}

// One block: out = {plaq, E_raw, Q, Q(t = 0..Nt-1)}.
__global__ __launch_bounds__(kBlock) void measure_finish_kernel(const double *__restrict__ partials, Geom g, int cps,
                                                                double *__restrict__ out)
{
    const int nblk = 2 * g.n[3] * cps;
    double acc[2] = {0.0, 0.0};
    for (int b = threadIdx.x; b < nblk; b += kBlock) {
        acc[0] += partials[3 * b + 0];
        acc[1] += partials[3 * b + 1];
    }
    block_sum<2>(acc);
    const double qnorm = 1.0 / (256.0 * M_PI * M_PI);
    for (int t = threadIdx.x; t < g.n[3]; t += kBlock) {
        double q = 0.0;
        for (int p = 0; p < 2; p++)
            for (int c = 0; c < cps; c++) q += partials[3 * ((p * g.n[3] + t) * cps + c) + 2];
        out[3 + t] = q * qnorm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double q = 0.0;
        for (int t = 0; t < g.n[3]; t++) q += out[3 + t];
        out[0] = acc[0] / (18.0 * (double)g.v);
        out[1] = acc[1] / (double)g.v;
        out[2] = q;
    }
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

static inline int chunks_per_slab(const Geom &g) { return (g.vs_h + kBlock - 1) / kBlock; }

int measure_partials_count(const Geom &g) { return 3 * 2 * g.n[3] * chunks_per_slab(g); }
int plaquette_partials_count(const Geom &g) { return (g.v + kBlock - 1) / kBlock; }

template <class T> void launch_measure(const void *f, const Geom &g, double *partials, hipStream_t s)
{
    const int cps = chunks_per_slab(g);
    measure_kernel<T><<<2 * g.n[3] * cps, kBlock, 0, s>>>((const typename vec2<T>::type *)f, g, cps, partials);
}

void launch_measure_finish(const double *partials, const Geom &g, double *out, hipStream_t s)
{
    measure_finish_kernel<<<1, kBlock, 0, s>>>(partials, g, chunks_per_slab(g), out);
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
