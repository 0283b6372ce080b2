// glac_gpu.hpp -- GLAC's Action / Flow / Correlator / System interfaces for the
// hot path, implemented over the C ABI of libglacx.so (include/glacx.h).
//
// Header-only templates, so they instantiate against GLAC's own types: a
// lattice type L is anything with GLAC's public Lattice<SU3> shape
// (GLAC/math/lattice.h:23-29): `std::vector<SU3> m_sites`,
// `std::vector<unsigned int> m_dim`, `unsigned long m_latticeSize`, where
// SU3 is 18 doubles (`double mat[18]`, GLAC/math/matrices/su3.h:22).  A
// gauge field is GLAC's `L lattice[4]` (GLAC/system.cpp:158-161).
//
// Residency.  GLAC calls Flow::flowField(lattice) and then
// Correlator::calculate(lattice) every flow step (GLAC/system.cpp:494-508).
// The Device keeps each bound host field resident on the GPU and tracks which
// copy is newer, so such a loop moves no lattice data between host and device:
//   - the first use of a host field uploads it (or after markHostModified);
//   - device-side updates (flow, Metropolis) mark the device copy newer;
//   - syncHost() downloads a device-newer field before host code reads it
//     (e.g. before GLAC's IO::FieldIO::writeFieldToFile).
//
// Errors follow GLAC's convention of a fatal, printed exit
// (Parallel::Communicator::MPIExit, GLAC/parallelization/communicator.cpp:
// 526-533), except that the exit status is non-zero: GLAC exits with 0.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <utility>
#include <vector>

#include "../../include/glacx.h"

namespace glac_gpu {

[[noreturn]] inline void fatal(const std::string &msg)
{
    std::fprintf(stderr, "glac_gpu: %s\n", msg.c_str());
    std::exit(EXIT_FAILURE);
}

// ---------------------------------------------------------------- device ---
class Device {
public:
    Device(const std::vector<unsigned int> &dims, int precision = GLACX_FP64, int device = 0)
    {
        if (dims.size() != 4) fatal("lattice needs 4 extents");
        for (int k = 0; k < 4; k++) m_dims[k] = (int)dims[k];
        m_volume = (unsigned long)m_dims[0] * m_dims[1] * m_dims[2] * m_dims[3];
        const int rc = glacx_create(m_dims, precision, device, &m_ctx);
        if (rc != GLACX_OK)
            fatal(std::string("glacx_create failed: ") + (m_ctx ? glacx_last_error(m_ctx) : "bad lattice extents"));
    }
    ~Device() { glacx_destroy(m_ctx); }
    Device(const Device &) = delete;
    Device &operator=(const Device &) = delete;

    glacx_ctx *ctx() { return m_ctx; }
    const int *dims() const { return m_dims; }
    unsigned long volume() const { return m_volume; }

    void check(int rc, const char *what)
    {
        if (rc != GLACX_OK) fatal(std::string(what) + ": " + glacx_last_error(m_ctx));
    }

    // Device field holding `lattice` (uploading it if needed).
    template <class L> int acquire(L *lattice)
    {
        Binding &b = bind(lattice);
        if (b.host_newer) {
            upload(lattice, b.field);
            b.host_newer = false;
            b.device_newer = false;
        }
        return b.field;
    }
    // The device copy of `lattice` was just modified on the GPU.
    template <class L> void markDeviceModified(L *lattice) { bind(lattice).device_newer = true; }
    // Host code modified `lattice`: the next device use uploads it.
    template <class L> void markHostModified(L *lattice)
    {
        Binding &b = bind(lattice);
        b.host_newer = true;
        b.device_newer = false;
    }
    // Bring the host copy up to date.
    template <class L> void syncHost(L *lattice)
    {
        Binding &b = bind(lattice);
        if (b.device_newer) {
            double *links[4];
            for (int mu = 0; mu < 4; mu++) links[mu] = sites(lattice[mu]);
            check(glacx_download_field(m_ctx, b.field, links), "download");
            b.device_newer = false;
        }
    }

    // Device field that `lattice` is about to be overwritten into on the GPU
    // (a start or a file load): nothing is uploaded, the device copy is newer.
    template <class L> int overwrite(L *lattice)
    {
        Binding &b = bind(lattice);
        b.host_newer = false;
        b.device_newer = true;
        return b.field;
    }
    // Give `lattice` a device field now (its host copy is uploaded on first use).
    template <class L> void reserve(L *lattice) { bind(lattice); }
    // dst <- src on the device (System::copyToFlowLattice, GLAC/system.cpp:
    // 179-188); both lattices must hold device fields of their own.
    template <class L> void copyField(L *dst, L *src)
    {
        const int fs = acquire(src);
        Binding &bd = bind(dst);
        if (bd.field == fs || !bound(src)) fatal("copyField: source and destination need device fields of their own");
        check(glacx_copy_field(m_ctx, bd.field, fs), "copy_field");
        bd.host_newer = false;
        bd.device_newer = true;
    }
    // Starts of System::latticeSetup (GLAC/system.cpp:200-224) on the device.
    template <class L> void hotStart(L *lattice, unsigned long long seed, int kind, double eps)
    {
        check(glacx_hot_start(m_ctx, overwrite(lattice), seed, kind, eps), "hot_start");
    }
    template <class L> void coldStart(L *lattice) { check(glacx_cold_start(m_ctx, overwrite(lattice)), "cold_start"); }

    // Read a configuration file straight into the device copy of `lattice`
    // (GLAC's .bin, or Chroma's big-endian variant); the host copy is then
    // stale until syncHost().
    template <class L> void loadInto(L *lattice, const std::string &path, bool chroma)
    {
        Binding &b = bind(lattice);
        check(chroma ? glacx_load_chroma(m_ctx, b.field, path.c_str()) : glacx_load_bin(m_ctx, b.field, path.c_str()),
              chroma ? "load_chroma" : "load_bin");
        b.host_newer = false;
        b.device_newer = true;
    }

private:
    struct Binding {
        const void *host = nullptr;
        int field = 0;
        bool host_newer = true;
        bool device_newer = false;
    };

    template <class Lmu> double *sites(Lmu &l)
    {
        if (l.m_latticeSize != m_volume || l.m_sites.size() != m_volume) fatal("lattice size does not match the device");
        static_assert(sizeof(l.m_sites[0]) == 18 * sizeof(double), "SU3 must be 18 doubles");
        return reinterpret_cast<double *>(&l.m_sites[0]);
    }

    template <class L> bool bound(L *lattice) const
    {
        for (const Binding &b : m_bind)
            if (b.host == lattice) return true;
        return false;
    }

    template <class L> Binding &bind(L *lattice)
    {
        for (Binding &b : m_bind)
            if (b.host == lattice) return b;
        if (m_bind.size() == 2) {
            // evict the older binding, writing it back first
            Binding &old = m_bind.front();
            if (old.device_newer) fatal("two device fields are in use and the evicted one is newer on the device; "
                                        "call syncHost() first");
            m_bind.erase(m_bind.begin());
        }
        Binding b;
        b.host = lattice;
        b.field = m_bind.empty() ? GLACX_FIELD_MAIN : (m_bind[0].field == GLACX_FIELD_MAIN ? GLACX_FIELD_FLOW
                                                                                             : GLACX_FIELD_MAIN);
        m_bind.push_back(b);
        return m_bind.back();
    }

    template <class L> void upload(L *lattice, int field)
    {
        const double *links[4];
        for (int mu = 0; mu < 4; mu++) links[mu] = sites(lattice[mu]);
        check(glacx_upload_field(m_ctx, field, links), "upload");
    }

    glacx_ctx *m_ctx = nullptr;
    int m_dims[4];
    unsigned long m_volume = 0;
    std::vector<Binding> m_bind;
};

// ------------------------------------------------------------ 3x3 helpers ---
// GLAC's SU3 storage: U_ij at mat[6i + 2j] (re), mat[6i + 2j + 1] (im).
inline void mul(const double *a, const double *b, double *c)
{
    double t[18];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double re = 0, im = 0;
            for (int k = 0; k < 3; k++) {
                re += a[6 * i + 2 * k] * b[6 * k + 2 * j] - a[6 * i + 2 * k + 1] * b[6 * k + 2 * j + 1];
                im += a[6 * i + 2 * k] * b[6 * k + 2 * j + 1] + a[6 * i + 2 * k + 1] * b[6 * k + 2 * j];
            }
            t[6 * i + 2 * j] = re;
            t[6 * i + 2 * j + 1] = im;
        }
    for (int n = 0; n < 18; n++) c[n] = t[n];
}
inline void dag(const double *a, double *c)
{
    double t[18];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            t[6 * i + 2 * j] = a[6 * j + 2 * i];
            t[6 * i + 2 * j + 1] = -a[6 * j + 2 * i + 1];
        }
    for (int n = 0; n < 18; n++) c[n] = t[n];
}

// ------------------------------------------------------------------ action ---
// WilsonGaugeAction (GLAC/actions/action.h:49-68, wilsongaugeaction.cpp:6-86).
// computeStaple / getDeltaAction keep GLAC's per-link FP64 host semantics (the
// staple is cached between the two calls); getActionDerivative runs on the
// GPU.  A GPU sweep replaces the per-link calls: see Metropolis below.
template <class SU3, class L> class WilsonGaugeAction {
public:
    WilsonGaugeAction(double beta, Device *device) : m_beta(beta), m_dev(device) {}

    double getDeltaAction(SU3 &U, SU3 &UPrime)
    {
        // -beta/3 Re tr[(U' - U) A]
        double s = 0.0;
        for (int i = 0; i < 3; i++)
            for (int k = 0; k < 3; k++) {
                const double dr = UPrime.mat[6 * i + 2 * k] - U.mat[6 * i + 2 * k];
                const double di = UPrime.mat[6 * i + 2 * k + 1] - U.mat[6 * i + 2 * k + 1];
                s += dr * m_staple[6 * k + 2 * i] - di * m_staple[6 * k + 2 * i + 1];
            }
        return -m_beta / 3.0 * s;
    }

    void computeStaple(L *lattice, int i, int j, int k, int l, int mu)
    {
        const std::vector<unsigned int> &N = lattice[0].m_dim;
        const int c[4] = {i, j, k, l};
        auto site = [&](int d1, int s1, int d2, int s2) {
            int m[4] = {c[0], c[1], c[2], c[3]};
            m[d1] = (m[d1] + s1 + (int)N[d1]) % (int)N[d1];
            if (s2) m[d2] = (m[d2] + s2 + (int)N[d2]) % (int)N[d2];
            return (unsigned long)m[0] + N[0] * ((unsigned long)m[1] + N[1] * ((unsigned long)m[2] + N[2] * m[3]));
        };
        auto link = [&](int dir, unsigned long s) { return lattice[dir].m_sites[s].mat; };
        for (int n = 0; n < 18; n++) m_staple[n] = 0.0;
        for (int nu = 0; nu < 4; nu++) {
            if (nu == mu) continue;
            double t1[18], t2[18], d[18];
            // U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
            dag(link(mu, site(nu, 1, mu, 0)), d);
            mul(link(nu, site(mu, 1, nu, 0)), d, t1);
            dag(link(nu, site(0, 0, 0, 0)), d);
            mul(t1, d, t1);
            // U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu)
            dag(link(nu, site(mu, 1, nu, -1)), t2);
            dag(link(mu, site(nu, -1, mu, 0)), d);
            mul(t2, d, t2);
            mul(t2, link(nu, site(nu, -1, mu, 0)), t2);
            for (int n = 0; n < 18; n++) m_staple[n] += t1[n] + t2[n];
        }
    }

    // Whole-lattice flow force Z_mu on the GPU.
    L getActionDerivative(L *lattice, int mu)
    {
        const int f = m_dev->acquire(lattice);
        L out(lattice[mu].m_dim);
        m_dev->check(glacx_action_derivative(m_dev->ctx(), f, mu, reinterpret_cast<double *>(&out.m_sites[0])),
                     "action_derivative");
        return out;
    }

    const double *staple() const { return m_staple; }

private:
    double m_beta;
    Device *m_dev;
    double m_staple[18];
};

// -------------------------------------------------------------------- flow ---
// Flow (GLAC/flow/flow.h:40-42): one RK3 step, in place.
template <class L> class Flow {
public:
    Flow(Device *device, double epsilon) : m_dev(device), m_eps(epsilon) {}
    void flowField(L *lattice)
    {
        const int f = m_dev->acquire(lattice);
        m_dev->check(glacx_flow(m_dev->ctx(), f, 1, m_eps, nullptr, nullptr, nullptr, nullptr), "flow");
        m_dev->markDeviceModified(lattice);
    }

private:
    Device *m_dev;
    double m_eps;
};

// ------------------------------------------------------------- observables ---
// MasterSamplerTopcXYZ (GLAC/observables/mastersamplertopcxyz.cpp:236-384):
// per observation index the plaquette, topological charge, E_raw and the
// per-time-slice charge.  getObservablesVector / copyObservable keep GLAC's
// ordering {plaq, topc, energy, topct[0..Nt-1]} (:227-244).
template <class L> class MasterSamplerTopcXYZ {
public:
    MasterSamplerTopcXYZ(Device *device, unsigned int nObs)
        : m_dev(device), m_nt(device->dims()[3]), m_obs(nObs, std::vector<double>(3 + device->dims()[3], 0.0))
    {
    }

    void calculate(L *lattice, unsigned int iObs)
    {
        if (iObs >= m_obs.size()) fatal("observation index out of range");
        const int f = m_dev->acquire(lattice);
        std::vector<double> &o = m_obs[iObs];
        m_dev->check(glacx_measure(m_dev->ctx(), f, &o[0], &o[2], &o[1], &o[3]), "measure");
    }

    // GLAC's System::flowConfiguration loop (GLAC/system.cpp:494-508) in one
    // call: calculate(lattice, 0), then nFlows times {flowField; calculate},
    // the lattice flowed in place and observations 0..nFlows filled.  Each
    // step's first RK3 stage measures the field it reads, which saves the
    // separate clover pass of the step-by-step loop.
    void flowConfiguration(L *lattice, unsigned int nFlows, double epsilon)
    {
        if (nFlows + 1 > m_obs.size()) fatal("flowConfiguration: more flow steps than observations");
        const int f = m_dev->acquire(lattice);
        const size_t m = nFlows + 1;
        std::vector<double> p(m), e(m), q(m), qt(m * (size_t)m_nt);
        m_dev->check(glacx_flow(m_dev->ctx(), f, (int)nFlows, epsilon, p.data(), e.data(), q.data(), qt.data()),
                     "flow");
        m_dev->markDeviceModified(lattice);
        for (size_t i = 0; i < m; i++) {
            std::vector<double> &o = m_obs[i];
            o[0] = p[i];
            o[1] = q[i];
            o[2] = e[i];
            std::copy(qt.begin() + i * m_nt, qt.begin() + (i + 1) * m_nt, o.begin() + 3);
        }
    }

    std::vector<double> getObservablesVector(unsigned int iObs) const { return m_obs[iObs]; }
    void copyObservable(unsigned int iObs, const std::vector<double> &obs) { m_obs[iObs] = obs; }
    double getObservable(unsigned int iObs) const { return m_obs[iObs][0]; }
    void reset()
    {
        for (auto &o : m_obs) std::fill(o.begin(), o.end(), 0.0);
    }

private:
    Device *m_dev;
    int m_nt;
    std::vector<std::vector<double>> m_obs;
};

// ----------------------------------------------------------------- field IO ---
// IO::FieldIO (GLAC/io/fieldio.cpp:61-122, 195-248, 259-338) on device-resident
// lattices: loads go to the device copy, writes read the device copy, in
// GLAC's .bin layout (site-major, 4 links per site, 18 doubles per link).
// Paths are complete (GLAC prefixes Parameters::getFilePath() and the
// input / output folders itself).
template <class L> class FieldIO {
public:
    explicit FieldIO(Device *device) : m_dev(device) {}
    void loadFieldConfiguration(const std::string &path, L *lattice) { m_dev->loadInto(lattice, path, false); }
    void loadChromaFieldConfiguration(const std::string &path, L *lattice) { m_dev->loadInto(lattice, path, true); }
    void writeFieldToFile(const std::string &path, L *lattice)
    {
        const int f = m_dev->acquire(lattice);
        m_dev->check(glacx_save_bin(m_dev->ctx(), f, path.c_str()), "save_bin");
    }

private:
    Device *m_dev;
};

// --------------------------------------------------------------- Metropolis ---
// System::update (GLAC/system.cpp:333-361) as one GPU checkerboard sweep.
// The acceptance rate of each sweep is accumulated as GLAC's m_acceptanceScore
// (:359-360).
template <class L> class Metropolis {
public:
    Metropolis(Device *device, double beta, double eps, int nUpdates, unsigned long long seed, int reunitEvery = 1)
        : m_dev(device), m_beta(beta), m_eps(eps), m_nhit(nUpdates), m_seed(seed), m_reunit(reunitEvery)
    {
    }
    // nSweeps sweeps (GLAC calls update() once per sweep); each sweep's
    // acceptance rate is added to the score.
    void update(L *lattice, int nSweeps = 1)
    {
        const int f = m_dev->acquire(lattice);
        if (f != GLACX_FIELD_MAIN) fatal("the Markov chain must live in the main device field");
        std::vector<double> acc((size_t)nSweeps);
        m_dev->check(glacx_metropolis(m_dev->ctx(), nSweeps, m_beta, m_eps, m_nhit, m_reunit, m_seed, m_sweep,
                                      acc.data(), nullptr),
                     "metropolis");
        m_sweep += (unsigned long long)nSweeps;
        for (double a : acc) m_acceptanceScore += a;
        m_dev->markDeviceModified(lattice);
    }
    void resetAcceptanceScore() { m_acceptanceScore = 0.0; }
    double acceptanceScore() const { return m_acceptanceScore; }
    unsigned long long sweeps() const { return m_sweep; }

private:
    Device *m_dev;
    double m_beta, m_eps;
    int m_nhit;
    unsigned long long m_seed, m_sweep = 0;
    int m_reunit;
    double m_acceptanceScore = 0.0;
};

// ------------------------------------------------ GLAC class hierarchy shims ---
// GLAC's System holds `Action *m_S` and `Correlator *m_correlator,
// *m_flowCorrelator`, chosen by the "action" and observable strings
// (GLAC/system.cpp:46-125).  These templates derive from a GLAC class given as
// `Base` and override its virtuals with the GPU path, so an instance can be
// stored in those members and called through them.

// Base: GLAC's Action (GLAC/actions/action.h:22-69) or WilsonGaugeAction.
// computeStaple / getDeltaAction keep GLAC's per-link FP64 semantics on the
// host (the staple is cached between the two calls, as GLAC's own);
// getActionDerivative is the GPU flow force.
template <class Base, class SU3, class L> class GpuAction : public Base {
public:
    template <class... A>
    GpuAction(double beta, Device *device, A &&...baseArgs) : Base(std::forward<A>(baseArgs)...), m_impl(beta, device)
    {
    }
    double getDeltaAction(SU3 &U, SU3 &UPrime) override { return m_impl.getDeltaAction(U, UPrime); }
    void computeStaple(L *lattice, const int i, const int j, const int k, const int l, const int mu) override
    {
        m_impl.computeStaple(lattice, i, j, k, l, mu);
    }
    L getActionDerivative(L *lattice, const int mu) override { return m_impl.getActionDerivative(lattice, mu); }

private:
    WilsonGaugeAction<SU3, L> m_impl;
};

// Base: GLAC's MasterSamplerTopcXYZ (GLAC/observables/mastersamplertopcxyz.h)
// or any Correlator whose copyObservable(iObs, obs) stores GLAC's observation
// vector {plaq, topc, energy, topct[0..Nt-1]} (mastersamplertopcxyz.cpp:
// 213-235).  calculate() measures on the GPU and hands the vector to the
// base, so GLAC's own storers, printing, statistics and file writers see it.
template <class Base, class L> class GpuObservables : public Base {
public:
    template <class... A>
    GpuObservables(Device *device, A &&...baseArgs) : Base(std::forward<A>(baseArgs)...), m_dev(device)
    {
    }
    void calculate(L *lattice, const unsigned int iObs) override
    {
        const int f = m_dev->acquire(lattice);
        std::vector<double> o(3 + (size_t)m_dev->dims()[3]);
        m_dev->check(glacx_measure(m_dev->ctx(), f, &o[0], &o[2], &o[1], &o[3]), "measure");
        Base::copyObservable(iObs, o);
    }
    // GLAC's System::flowConfiguration loop (GLAC/system.cpp:494-508) in one
    // call: observations 0..nFlows of the lattice flowed in place, each step's
    // first RK3 stage measuring the field it reads.
    void flowConfiguration(L *lattice, unsigned int nFlows, double epsilon)
    {
        const int f = m_dev->acquire(lattice);
        const int nt = m_dev->dims()[3];
        const size_t m = nFlows + 1;
        std::vector<double> p(m), e(m), q(m), qt(m * (size_t)nt);
        m_dev->check(glacx_flow(m_dev->ctx(), f, (int)nFlows, epsilon, p.data(), e.data(), q.data(), qt.data()),
                     "flow");
        m_dev->markDeviceModified(lattice);
        std::vector<double> o(3 + (size_t)nt);
        for (size_t i = 0; i < m; i++) {
            o[0] = p[i];
            o[1] = q[i];
            o[2] = e[i];
            std::copy(qt.begin() + i * nt, qt.begin() + (i + 1) * nt, o.begin() + 3);
            Base::copyObservable((unsigned int)i, o);
        }
    }

private:
    Device *m_dev;
};

// --------------------------------------------------------------- run loop ---
// The keys of GLAC's JSON configuration the path consumes
// (GLAC/config/configloader.cpp:55-181, defaults GLAC/config/parameters.cpp:
// 5-93).
struct RunParameters {
    std::vector<unsigned int> N{8, 8, 8, 16};  // NSpatial x 3, NTemporal
    double beta = 6.0, SU3Eps = 0.24, flowEpsilon = 0.01;
    unsigned int NTherm = 10, NCf = 30, NCor = 10, NUpdates = 10, NFlows = 0;
    bool hotStart = true, RSTHotStart = false;
    bool storeThermalizationObservables = false;
    unsigned long long metropolisSeed = 1, randomMatrixSeed = 2;
    int precision = GLACX_FP64;
    int reunitEvery = 1;      // FP64 reunitarisation every this many sweeps (0: never, as GLAC)
    int device = 0;
    std::string configPrefix;  // non-empty: write every configuration as <prefix>_config%05d.bin
    bool verbose = false;
};

// GLAC's System (GLAC/system.h:117-121): latticeSetup() then run(), i.e.
// thermalize + runMetropolis with a flowConfiguration per configuration
// (GLAC/system.cpp:193-231, 262-319, 363-517), on the device.  Corr and
// FlowCorr are GLAC Correlator types (typically GpuObservables over GLAC's
// MasterSamplerTopcXYZ, measuring on the same Device); the flow correlator
// also needs flowConfiguration().  The Device (built from p.N, p.precision,
// p.device) holds the chain lattice in its main field and the flow lattice in
// the other, GLAC's Lattice<SU3>[4] each; syncHost() brings the chain to the
// host.
template <class L, class Corr, class FlowCorr> class System {
public:
    System(Device *device, const RunParameters &p, Corr *correlator, FlowCorr *flowCorrelator)
        : m_p(p), m_dev(*device), m_correlator(correlator), m_flowCorrelator(flowCorrelator),
          m_update(device, p.beta, p.SU3Eps, (int)p.NUpdates, p.metropolisSeed, p.reunitEvery)
    {
        for (int k = 0; k < 4; k++)
            if ((int)p.N.at(k) != device->dims()[k]) fatal("System: lattice extents differ from the device's");
        for (int mu = 0; mu < 4; mu++) {
            m_lattice[mu] = L(p.N);
            m_flowLattice[mu] = L(p.N);
        }
    }

    Device &device() { return m_dev; }
    L *lattice() { return m_lattice; }
    L *flowLattice() { return m_flowLattice; }
    double getAcceptanceRate() const { return m_acceptanceRate; }

    // Hot start (generateRST if RSTHotStart, else GLAC's generateRandom,
    // det = -1) or cold start, GLAC/system.cpp:200-224, from the device's
    // link streams of randomMatrixSeed.  The chain lattice takes the main
    // device field, the flow lattice the other one.
    void latticeSetup()
    {
        if (m_p.hotStart)
            m_dev.hotStart(m_lattice, m_p.randomMatrixSeed, m_p.RSTHotStart ? GLACX_HOT_RST : GLACX_HOT_GLAC_RANDOM,
                           m_p.SU3Eps);
        else
            m_dev.coldStart(m_lattice);
        m_dev.reserve(m_flowLattice);
    }

    void run()
    {
        thermalize();
        runMetropolis();
    }

    void thermalize()
    {
        m_NThermSteps = m_p.storeThermalizationObservables ? 1 + m_p.NTherm : 0;
        if (m_p.storeThermalizationObservables) {
            m_correlator->calculate(m_lattice, 0);
            for (unsigned int i = 1; i <= m_p.NTherm; i++) {
                m_update.update(m_lattice);
                m_correlator->calculate(m_lattice, i);
            }
        } else if (m_p.NTherm > 0) {
            m_update.update(m_lattice, (int)m_p.NTherm);
        }
        if (m_p.verbose && m_p.NTherm)
            std::printf("Termalization complete. Acceptance rate: %f\n", m_update.acceptanceScore() / m_p.NTherm);
    }

    void runMetropolis()
    {
        m_update.resetAcceptanceScore();
        for (unsigned int iConfig = 0; iConfig < m_p.NCf; iConfig++) {
            m_update.update(m_lattice, (int)m_p.NCor);
            if (m_p.NFlows != 0) {
                m_dev.copyField(m_flowLattice, m_lattice);
                flowConfiguration(iConfig);
                m_correlator->copyObservable(iConfig + m_NThermSteps, m_flowCorrelator->getObservablesVector(0));
            } else {
                m_correlator->calculate(m_lattice, iConfig + m_NThermSteps);
            }
            if (m_p.verbose) {
                std::printf("%-4u ", iConfig);
                m_correlator->printObservable(iConfig + m_NThermSteps);
                std::printf("\n");
            }
            if (!m_p.configPrefix.empty()) {
                char name[32];
                std::snprintf(name, sizeof name, "_config%05u.bin", iConfig);
                m_dev.check(glacx_save_bin(m_dev.ctx(), m_dev.acquire(m_lattice), (m_p.configPrefix + name).c_str()),
                            "save_bin");
            }
        }
        m_acceptanceRate = m_p.NCf ? m_update.acceptanceScore() / double(m_p.NCf * m_p.NCor) : 0.0;
        m_correlator->runStatistics();
        m_correlator->writeObservableToFile(m_acceptanceRate);
    }

    // System::flowConfiguration (GLAC/system.cpp:471-517): reset, observations
    // 0..NFlows of the flowed copy, then the flow files of configuration iConfig.
    void flowConfiguration(unsigned int iConfig)
    {
        m_flowCorrelator->reset();
        m_flowCorrelator->flowConfiguration(m_flowLattice, m_p.NFlows, m_p.flowEpsilon);
        m_flowCorrelator->writeFlowObservablesToFile(iConfig);
    }

    void syncHost() { m_dev.syncHost(m_lattice); }

private:
    RunParameters m_p;
    Device &m_dev;
    Corr *m_correlator;
    FlowCorr *m_flowCorrelator;
    Metropolis<L> m_update;
    L m_lattice[4], m_flowLattice[4];
    unsigned int m_NThermSteps = 0;
    double m_acceptanceRate = 0.0;
};

} // namespace glac_gpu
