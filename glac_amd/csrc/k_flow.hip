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
// FP32 links are reunitarised (FP64 Gram-Schmidt) as the last stage of a
// step stores them.
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

// ---- clover observables fused into stage 0 -------------------------------
// Stage 0 of a flow step reads the field the previous step produced, which is
// the field GLAC measures between steps (GLAC/system.cpp:497-503).  Its staple
// terms already hold three of the four clover leaves at x: for wavefront mu
// and term nu, with
//   up  = U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag,
//   low = U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu),
// U_mu(x) up is the leaf L1 of plane (mu, nu), U_mu(x) low its leaf L2 and
// also L4^dag of plane (nu, mu) (GLAC's leaves, GLAC/observables/
// mastersamplertopcxyz.cpp:255-328).  The clover C = L1 - L2 + L3 - L4 enters
// the observables only through F = traceless(C - C^dag) (:331-340), which is
// linear in the leaves; with T(M) = traceless(M - M^dag) and T(-M) = T(M^dag),
//   F = T(U_mu(x) (up - low)) + T(L3) + T(U_nu(x) low'),
// low' being wavefront nu's lower term for mu.  So for each of GLAC's planes
// (0,1) (0,2) (0,3) (2,3) (3,1) (1,2) the first-index wavefront adds
// T(U_mu (up - low)), the second-index one T(U_nu low'), and one of the two
// also T(L3), L3 = U_mu(x-mu)^dag U_nu(x-mu-nu)^dag U_mu(x-mu-nu) U_nu(x-nu)
// (two more gathers; both backward links are in LDS).  Every wavefront forms
// one extra product per term, and the L3 leaves are dealt 2, 2, 1, 1.
//
// F of a plane is 8 reals {Im F00, Im F11, F01, F02, F12} in one LDS slot
// per plane.  The three staple terms run as three phases separated by
// barriers, in a per-wavefront order chosen so that the two wavefronts of a
// plane reach it in different phases: the earlier one stores its
// contribution, the later one adds its own to it.  No atomics, and the sum
// of the two is formed in a fixed order.  The L3 leaves go two per phase.
// After the last phase wavefront w < 3 evaluates the pair (w, w + 3) =
// ((0, w+1), (rho, sigma)) per site as the clover kernel does:
//   E += Re tr(F_a F_a) + Re tr(F_b F_b),  q -= Re tr(F_a F_b),
// and the plaquette sum takes Re tr L1 of each plane from its first-index
// wavefront.  The block's sums go to the clover kernel's (chunk, t) slot.

// With FP64 links the F slots (24.6 KB per block) and the eight shared links
// (73.7 KB) would leave room for one block per CU; the FP64 measuring stage
// therefore shares only U_mu(x) through LDS and gathers the backward links
// U_nu(x - nu) from memory (four more gathers per link), which keeps it at
// two blocks per CU.

// Slot of GLAC's plane {d1, d2} stored as (d1, d2): 0..2 = (0, nu), 3 = (2,3),
// 4 = (3,1), 5 = (1,2); (mu, nu) is stored as such when mu = 0 or nu = mu % 3 + 1.
__device__ __forceinline__ bool plane_first(int mu, int nu) { return mu == 0 || nu == mu % 3 + 1; }
__device__ __forceinline__ int plane_slot(int d1, int d2) { return d1 == 0 ? d2 - 1 : (d1 == 1 ? 5 : d1 + 1); }
// Phase schedule, entry e = 3 mu + k for wavefront mu in phase k:
//   kPhaseNu:  the term nu (4 bits each): mu 0: 1 2 3, mu 1: 2 0 3, mu 2: 0 3 1, mu 3: 0 1 2
//   kPhaseAdd: the plane's other wavefront stored first (add to the slot)
//   kPhaseL3:  this wavefront also adds the plane's L3 leaf
constexpr unsigned long long kPhaseNu = 0x210130302321ull;
constexpr unsigned kPhaseAdd = 0x936u;  // entries 1 2 4 5 8 11
constexpr unsigned kPhaseL3 = 0x3a3u;   // entries 0 1 5 7 8 9

// SU(3) product from the first two rows of `a` (all that load_rows01 reads):
// the product's first two rows, then its third rebuilt from them.
template <bool DB, class T> __device__ __forceinline__ M3<T> su3_product(const M3<T> &a, const M3<T> &b)
{
    M3<T> p;
    mm<false, DB, false, 2>(p, a, b);
    rebuild_third_row(p);
    return p;
}

// Order in which a launch visits the 32-site chunks of each parity.  In
// storage order a slab's chunks run (lt, lz, y-chunk), so a chunk's
// t-neighbours come b_z x (plane) chunks later, beyond an XCD's L2 on large
// lattices.  When every x-y plane holds whole chunks (Ny Nx / 2 divisible by
// 32) the launch instead runs (y-chunk, lz, lt) within each slab and parity:
// z- and t-neighbours inside the slab are then at most b_z b_t chunks apart.
__device__ __forceinline__ int flow_chunk(const Geom &g, int b)
{
    const int plane = g.run >> g.lbz;  // Ny Nx / 2 entries per x-y plane of one parity
    if (plane & 31) return b;
    const int cpp = plane >> 5, per = g.bh >> 5;  // chunks per plane, per slab-parity
    const int sp = fdiv(b << 5, g.d_bh), r = b - sp * per;
    const int lt = r & ((1 << g.lbt) - 1), lz = (r >> g.lbt) & ((1 << g.lbz) - 1), yc = r >> (g.lbt + g.lbz);
    return sp * per + ((lt << g.lbz) + lz) * cpp + yc;
}

template <class T> struct FlowMeasLds {
    T f[6][8][kFlowSites];
};
template <class T> __device__ __forceinline__ FlowMeasLds<T> &flow_meas_lds()
{
    __shared__ FlowMeasLds<T> s;
    return s;
}

// FP64 tuning knobs (build time): waves per SIMD the stage is compiled for,
// and whether its plain stages share the backward links through LDS.
#ifndef GLACX_FP64_FLOW_WAVES
#define GLACX_FP64_FLOW_WAVES 2
#endif
#ifndef GLACX_FP64_BACK_LDS
#define GLACX_FP64_BACK_LDS 1
#endif

template <class T, int R, int STAGE, bool MEAS>
__global__ __launch_bounds__(kBlock, sizeof(T) == 4 ? (MEAS ? 3 : 4) : (MEAS ? 2 : GLACX_FP64_FLOW_WAVES)) void flow_stage_kernel(const typename vec2<T>::type *__restrict__ win,
                                                            typename vec2<T>::type *__restrict__ wout,
                                                            T *__restrict__ acc, Geom g, T cz, T ca,
                                                            double *__restrict__ meas)
{
    static_assert(!MEAS || STAGE == 0, "the measurement is fused into stage 0 only");
    // backward links U_nu(x - nu) shared through LDS, except in the FP64 measuring stage
    constexpr bool BACK_LDS = !(sizeof(T) == 8 && (MEAS || !GLACX_FP64_BACK_LDS));
    __shared__ LdsMats<T, BACK_LDS ? 8 : 4> sh;
    const int mu = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // one direction per wavefront
    const int lane = threadIdx.x & 63;
    const int hb = flow_chunk(g, xcd_block(blockIdx.x, gridDim.x));
    const int h_raw = hb * (kFlowSites / 2) + (lane & 31);
    const bool valid = h_raw < g.vh;
    const int h = valid ? h_raw : g.vh - 1;  // idle lanes of a partial block mirror a real site, store nothing
    const int idx = parity_index(g, lane >> 5, h);
    const Site x = site_of(g, idx);
    // stages 1 and 2 read the link's accumulator now, so that its latency
    // passes under the staple loop instead of stalling the step after it
    su3alg<T> cacc;
    if constexpr (STAGE != 0) cacc = acc_load(acc, g, mu, idx);
    const DirView<T> vmu = dir_view<T>(win, g, mu);
    // the site's neighbours x + mu and x - mu, shared with the other
    // wavefronts, whose staples step the same way in direction mu
    __shared__ int nbr[2][4][kFlowSites];
    const Site xpm = site_step(g, x, mu, +1);
    const int ipm = xpm.idx, imm = site_step(g, x, mu, -1).idx;
    nbr[0][mu][lane] = ipm;
    nbr[1][mu][lane] = imm;
    lds_put(sh, mu, lane, load_link<T, R>(vmu, idx));
    if constexpr (BACK_LDS) lds_put(sh, 4 + mu, lane, load_link<T, R>(vmu, imm));
    T plaq = T(0);
    __syncthreads();
    // U_d(x - d), from LDS or gathered (x - d from the wavefront of d)
    auto back_link = [&](int d) -> M3<T> {
        if constexpr (BACK_LDS)
            return lds_get(sh, 4 + d, lane);
        else
            return load_link<T, R>(dir_view<T>(win, g, d), nbr[1][d][lane]);
    };
    // staple A_mu(x) = sum_{nu != mu} [ U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
    //                                 + U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu) ]
    // (GLAC/actions/wilsongaugeaction.cpp:29-53)
    M3<T> a = zero3<T>();
    // measuring stage with three stored rows: T(U_mu(x) A_mu(x)) in compact form
    constexpr bool kCompactForce = MEAS && R == 3;
    T tz[8] = {T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0)};
    {
#pragma unroll 1
        for (int k = 0; k < 3; k++) {
            const int e = 3 * mu + k;
            const int nu = MEAS ? (int)((kPhaseNu >> (4 * e)) & 0xF) : (mu + k + 1) & 3;
            const DirView<T> vnu = dir_view<T>(win, g, nu);
            const int ipn = nbr[0][nu][lane], imn = nbr[1][nu][lane];
            const int ipmmn = site_step(g, xpm, nu, -1).idx;
            if constexpr (MEAS) {
                // the same staple terms, kept apart for the clover leaves
                M3<T> up, low;
                const bool first = plane_first(mu, nu);
                const int d1 = first ? mu : nu, d2 = first ? nu : mu;
                const int slot = plane_slot(d1, d2);
                T f[8];
                if constexpr (R == 2) {
                    // SU(3) links (checked at upload): products from two rows,
                    // third rows rebuilt
                    M3<T> p;  // first two rows of U_nu(x+mu) U_mu(x+nu)^dag
                    mm<false, true, false, 2>(p, load_rows01<T, R>(vnu, ipm), load_link<T, R>(vmu, ipn));
                    mm<false, true, false, 2>(up, p, lds_get(sh, nu, lane));
                    rebuild_third_row(up);
                    const M3<T> q = su3_product<false>(load_rows01<T, R>(vmu, imn), load_link<T, R>(vnu, ipmmn));
                    mm<true, false, false>(low, q, back_link(nu));
                    add_to(a, up);
                    add_to(a, low);
                    const M3<T> u = lds_get(sh, mu, lane);
                    if (first) {
                        plaq += retrace_mul(u, up);
                        sub_from(up, low);
                        ta2_product<false>(u, up, f);
                    } else {
                        ta2_product<false>(u, low, f);
                    }
                } else {
                    // three stored rows may hold GLAC's det = -1 links: full
                    // products.  The upper term is finished (its leaf's
                    // plaquette and T(U_mu up)) before the lower one is
                    // formed, which keeps fewer matrices live: T(U_mu (up -
                    // low)) = T(U_mu up) - T(U_mu low).  The force needs only
                    // T(U_mu A) = sum of T(U_mu up) + T(U_mu low) over the
                    // terms, so it is accumulated in that compact form (8
                    // reals instead of the 3 x 3 staple sum).
                    mm<false, true, false>(up, mul_ad(load_link<T, R>(vnu, ipm), load_link<T, R>(vmu, ipn)),
                                           lds_get(sh, nu, lane));
                    T tu[8];
                    {
                        const M3<T> u = lds_get(sh, mu, lane);
                        if (first) plaq += retrace_mul(u, up);
                        ta2_product<false>(u, up, tu);
                    }
                    mm<false, false, false>(low, mul_dd(load_link<T, R>(vnu, ipmmn), load_link<T, R>(vmu, imn)),
                                            back_link(nu));
                    T fl[8];
                    ta2_product<false>(lds_get(sh, mu, lane), low, fl);
#pragma unroll
                    for (int j = 0; j < 8; j++) {
                        tz[j] += tu[j] + fl[j];
                        f[j] = first ? tu[j] - fl[j] : fl[j];
                    }
                }
                if ((kPhaseL3 >> e) & 1) {
                    // L3 = (U_d2(y) U_d1(x-d1))^dag (U_d1(y) U_d2(x-d2)),  y = x - d1 - d2
                    const int iy = site_step(g, site_step(g, x, d1, -1), d2, -1).idx;
                    const DirView<T> v1 = first ? vmu : vnu, v2 = first ? vnu : vmu;
                    T f3[8];
                    if constexpr (R == 2) {
                        const M3<T> m1 = su3_product<false>(load_rows01<T, R>(v1, iy), back_link(d2));
                        const M3<T> m2 = su3_product<false>(load_rows01<T, R>(v2, iy), back_link(d1));
                        ta2_product<true>(m2, m1, f3);
                    } else {
                        // L3 as a chain U_d1(x-d1)^dag U_d2(y)^dag U_d1(y) U_d2(x-d2):
                        // the same three products with one matrix live between them
                        M3<T> c1, c2;
                        mm<true, true, false>(c1, back_link(d1), load_link<T, R>(v2, iy));
                        mm<false, false, false>(c2, c1, load_link<T, R>(v1, iy));
                        mm<false, false, false>(c1, c2, back_link(d2));
                        ta2_of(c1, f3);
                    }
#pragma unroll
                    for (int j = 0; j < 8; j++) f[j] += f3[j];
                }
                T(&fs)[8][kFlowSites] = flow_meas_lds<T>().f[slot];
                if ((kPhaseAdd >> e) & 1) {
#pragma unroll
                    for (int j = 0; j < 8; j++) f[j] = fs[j][lane] + f[j];
                }
#pragma unroll
                for (int j = 0; j < 8; j++) fs[j][lane] = f[j];
                __syncthreads();
            } else if constexpr (R == 2) {
                // Two stored rows: form each SU(3) product's first two rows
                // from operands of which only those are needed, and rebuild
                // the product's third row once (36 multiply-adds fewer per
                // product than rebuilding both operands).
                M3<T> p;  // U_nu(x+mu) U_mu(x+nu)^dag, first two rows
                mm<false, true, false, 2>(p, load_rows01<T, R>(vnu, ipm), load_link<T, R>(vmu, ipn));
                // the upper term p U_nu(x)^dag is SU(3) too: its first two rows
                // need only p's, and its third is rebuilt from them
                M3<T> t;
                mm<false, true, false, 2>(t, p, lds_get(sh, nu, lane));
                rebuild_third_row(t);
                add_to(a, t);
                M3<T> q;  // U_mu(x-nu) U_nu(x+mu-nu), whose dagger the lower staple uses
                mm<false, false, false, 2>(q, load_rows01<T, R>(vmu, imn), load_link<T, R>(vnu, ipmmn));
                rebuild_third_row(q);
                mm<true, false, true>(a, q, back_link(nu));
            } else {
                {
                    const M3<T> u1 = load_link<T, R>(vnu, ipm);
                    const M3<T> u2 = load_link<T, R>(vmu, ipn);
                    mm<false, true, true>(a, mul_ad(u1, u2), lds_get(sh, nu, lane));
                }
                {
                    const M3<T> u1 = load_link<T, R>(vnu, ipmmn);
                    const M3<T> u2 = load_link<T, R>(vmu, imn);
                    mm<false, false, true>(a, mul_dd(u1, u2), back_link(nu));
                }
            }
        }
    }
    if (valid) {
        const M3<T> u = lds_get(sh, mu, lane);
        su3alg<T> z;
        if constexpr (kCompactForce) {
            // Z = -1/2 T(U A) (the form force_from_omega evaluates)
            z.a0 = T(-0.5) * tz[0];
            z.a1 = T(-0.5) * tz[1];
#pragma unroll
            for (int k = 0; k < 6; k++) z.p[k] = T(-0.5) * tz[2 + k];
        } else {
            z = force_from_omega(mul(u, a));
        }
        su3alg<T> y;
        if (STAGE == 0) {
            acc_store(acc, g, mu, idx, z);
            y.a0 = cz * z.a0;
            y.a1 = cz * z.a1;
#pragma unroll
            for (int k = 0; k < 6; k++) y.p[k] = cz * z.p[k];
        } else {
            const su3alg<T> &c = cacc;
            y.a0 = cz * z.a0 - ca * c.a0;
            y.a1 = cz * z.a1 - ca * c.a1;
#pragma unroll
            for (int k = 0; k < 6; k++) y.p[k] = cz * z.p[k] - ca * c.p[k];
            if (STAGE == 1) acc_store(acc, g, mu, idx, y);
        }
        M3<T> w = zero3<T>();
        mm<false, false, false, R>(w, exp_su3(y), u);  // two rows suffice when two are stored
        if constexpr (sizeof(T) == 4 && STAGE == 2) {
            // FP32 links: project back onto the group in FP64 once per flow
            // step, as the last stage stores them.  Without it the rounding of
            // exp(Y) W drifts the links off it (|U U^dag - 1| ~ 1e-4 after 1000
            // steps at 24^3 x 48), and that drift, not the FP32 measurement,
            // dominates the error of E(t) at large t; within a step the two
            // unprojected stores add two roundings (~1e-7), which the
            // projection then removes.  Three stored rows are orthonormalised
            // row by row, which keeps det U (GLAC's det = -1 starts stay
            // det = -1); two stored rows are SU(3) and the third is rebuilt on
            // load.
            M3<double> d = convert<double>(w);
            if constexpr (R == 2)
                reunitarize<2, true>(d);
            else
                orthonormalize_rows<true>(d);
            w = convert<T>(d);
        }
        store_link<T, R>(wout, g, mu, idx, w);
    }
    if constexpr (MEAS) {
        // every lane is a real site here (the launch requires whole chunks);
        // the last phase's barrier has completed every slot
        double s[kMeasAcc] = {(double)plaq, 0.0, 0.0, 0.0};
        if (mu < 3) {
            T fa[8], fb[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                fa[k] = flow_meas_lds<T>().f[mu][k][lane];
                fb[k] = flow_meas_lds<T>().f[mu + 3][k][lane];
            }
            s[1] = retrace_ta2(fa, fa) + retrace_ta2(fb, fb);
            s[2] = -retrace_ta2(fa, fb);
        }
        block_sum<kMeasAcc>(s);
        if (threadIdx.x == 0) {
            // the block's 32 sites per parity are chunk jc of the run (blk, lt)
            const int h0 = hb * (kFlowSites / 2);
            const int blk = fdiv(h0, g.d_bh);
            const int rem = h0 - blk * g.bh;
            const int lt = fdiv(rem, g.d_run);
            const int jc = (rem - lt * g.run) >> 5;
            const int bt = fdiv(blk, g.d_nzb), zb = blk - bt * g.nzb;
            double *o = meas + kMeasAcc * ((zb * g.cpr + jc) * g.n[3] + (bt << g.lbt) + lt);
#pragma unroll
            for (int i = 0; i < kMeasAcc; i++) o[i] = s[i];
        }
    }
}

// Dynamic LDS reserved per block; nonzero values cap the blocks resident per
// CU (a tuning parameter of the launch, set at build time).
#ifndef GLACX_FLOW_LDS_BYTES
#define GLACX_FLOW_LDS_BYTES 0
#endif

// Fusing the measurement needs each block's 32 sites per parity to be one
// chunk of one run (a single time slice): run % 32 == 0, which holds for
// every lattice with bz Ny Nx / 2 divisible by 32.
bool flow_measure_fusable(const Geom &g) { return g.run % (kFlowSites / 2) == 0; }

template <class T, int R>
void launch_flow_stage(const void *win, void *wout, void *acc, const Geom &g, int stage, double eps, double *meas,
                       hipStream_t s)
{
    using V2 = typename vec2<T>::type;
    const unsigned blocks = (unsigned)((g.vh + kFlowSites / 2 - 1) / (kFlowSites / 2));
    const V2 *in = (const V2 *)win;
    V2 *out = (V2 *)wout;
    T *a = (T *)acc;
    switch (stage) {
    case 0:
        if (meas) {
            flow_stage_kernel<T, R, 0, true>
                <<<blocks, kBlock, GLACX_FLOW_LDS_BYTES, s>>>(in, out, a, g, T(0.25 * eps), T(0), meas);
            break;
        }
        flow_stage_kernel<T, R, 0, false>
            <<<blocks, kBlock, GLACX_FLOW_LDS_BYTES, s>>>(in, out, a, g, T(0.25 * eps), T(0), nullptr);
        break;
    case 1:
        flow_stage_kernel<T, R, 1, false><<<blocks, kBlock, GLACX_FLOW_LDS_BYTES, s>>>(
            in, out, a, g, T(8.0 / 9.0 * eps), T(17.0 / 36.0 * eps), nullptr);
        break;
    default:
        flow_stage_kernel<T, R, 2, false>
            <<<blocks, kBlock, GLACX_FLOW_LDS_BYTES, s>>>(in, out, a, g, T(0.75 * eps), T(1), nullptr);
        break;
    }
}

template void launch_flow_stage<float, 3>(const void *, void *, void *, const Geom &, int, double, double *,
                                         hipStream_t);
template void launch_flow_stage<float, 2>(const void *, void *, void *, const Geom &, int, double, double *,
                                         hipStream_t);
template void launch_flow_stage<double, 3>(const void *, void *, void *, const Geom &, int, double, double *,
                                          hipStream_t);

} // namespace glacx
