// Exercises glac_amd/host/glac_gpu.hpp the way GLAC's own code would, with a
// stand-in of GLAC's public Lattice<SU3> shape (GLAC/math/lattice.h:23-29).
//
//   host_adapter_check cpu   GLAC's dS = -36 test on a unit field
//                            (GLAC/tests/unit_tests/actiontests.cpp:67-114),
//                            directly and through GLAC's Action interface
//   host_adapter_check gpu [<dir>]   flowConfiguration-style loop (calculate,
//                            then flowField + calculate per step), the same
//                            loop as one flowConfiguration call (FP64, FP32)
//                            and Metropolis updates through the adapter,
//                            checked against the FP64 oracle, with the host
//                            copy left stale until syncHost; configurations
//                            written and read in <dir> through FieldIO (.bin
//                            and Chroma); then GLAC's System run loop
//                            (glac_gpu::System with GLAC-interface
//                            correlators) against the oracle's chain.
// Prints "PASS ..." lines; exits non-zero on the first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../glac_amd/host/glac_gpu.hpp"
#include "../oracle/glac_oracle.h"

struct SU3 {
    double mat[18];
};

template <class T> struct Lattice {
    std::vector<T> m_sites;
    std::vector<unsigned int> m_dim;
    unsigned long m_latticeSize = 0;
    Lattice() {}
    explicit Lattice(const std::vector<unsigned int> &dim) { allocate(dim); }
    void allocate(const std::vector<unsigned int> &dim)
    {
        m_dim = dim;
        m_latticeSize = (unsigned long)dim[0] * dim[1] * dim[2] * dim[3];
        m_sites.assign(m_latticeSize, T());
    }
};

// Stand-ins of GLAC's abstract bases with the signatures of
// GLAC/actions/action.h:49-68 and GLAC/observables/correlator.h:46-67 (GLAC's
// own headers need MPI), and of GLAC's MasterSamplerTopcXYZ storage
// (GLAC/observables/mastersamplertopcxyz.cpp:213-235): per observation the
// vector {plaq, topc, energy, topct[0..Nt-1]}.  The flow-file writer records
// the observations it would write.
class Action {
public:
    virtual ~Action() {}
    virtual double getDeltaAction(SU3 &U, SU3 &UPrime) = 0;
    virtual void computeStaple(Lattice<SU3> *lattice, const int i, const int j, const int k, const int l,
                               const int mu) = 0;
    virtual Lattice<SU3> getActionDerivative(Lattice<SU3> *lattice, const int mu) = 0;
};

class Correlator {
public:
    explicit Correlator(bool storeFlowObservable) : m_storeFlowObservable(storeFlowObservable) {}
    virtual ~Correlator() {}
    virtual void calculate(Lattice<SU3> *lattice, unsigned int iObs) = 0;
    virtual void writeObservableToFile(const double acceptanceRatio) { m_acceptanceWritten = acceptanceRatio; }
    virtual void writeFlowObservablesToFile(const unsigned int) {}
    virtual void runStatistics() {}
    virtual void printObservable(const unsigned int) {}
    virtual void printHeader() {}
    virtual void printStatistics() {}
    virtual void copyObservable(const unsigned int iObs, const std::vector<double> &obs) = 0;
    virtual std::vector<double> getObservablesVector(const unsigned int iObs) = 0;
    virtual double getObservable(const unsigned int iObs) { return getObservablesVector(iObs)[0]; }
    virtual void reset() {}
    bool m_storeFlowObservable;
    double m_acceptanceWritten = -1.0;
};

class MasterSamplerTopcXYZ : public Correlator {
public:
    MasterSamplerTopcXYZ(bool flow, unsigned int nObs, unsigned int nt)
        : Correlator(flow), m_obs(nObs, std::vector<double>(3 + nt, 0.0))
    {
    }
    void calculate(Lattice<SU3> *, unsigned int) override { std::abort(); }  // GLAC's CPU clover: not used here
    void copyObservable(const unsigned int iObs, const std::vector<double> &obs) override { m_obs.at(iObs) = obs; }
    std::vector<double> getObservablesVector(const unsigned int iObs) override { return m_obs.at(iObs); }
    void reset() override
    {
        for (auto &o : m_obs) std::fill(o.begin(), o.end(), 0.0);
    }
    void writeFlowObservablesToFile(const unsigned int iConfig) override
    {
        m_flowFiles.resize(iConfig + 1);
        m_flowFiles[iConfig] = m_obs;
    }
    std::vector<std::vector<double>> m_obs;
    std::vector<std::vector<std::vector<double>>> m_flowFiles;
};

#define REQUIRE(cond, ...)                                                                       \
    do {                                                                                          \
        if (!(cond)) {                                                                            \
            std::printf("FAIL line %d: ", __LINE__);                                             \
            std::printf(__VA_ARGS__);                                                             \
            std::printf("\n");                                                                    \
            return 1;                                                                             \
        }                                                                                         \
    } while (0)

static void set_unit(Lattice<SU3> *L)
{
    for (int mu = 0; mu < 4; mu++)
        for (auto &s : L[mu].m_sites) {
            for (double &v : s.mat) v = 0.0;
            s.mat[0] = s.mat[8] = s.mat[16] = 1.0;
        }
}

static int run_cpu()
{
    const std::vector<unsigned int> N = {4, 4, 4, 4};
    Lattice<SU3> L[4];
    for (int mu = 0; mu < 4; mu++) L[mu].allocate(N);
    set_unit(L);
    glac_gpu::WilsonGaugeAction<SU3, Lattice<SU3>> S(6.0, nullptr);
    SU3 two{};
    two.mat[0] = two.mat[8] = two.mat[16] = 2.0;
    for (unsigned x = 0; x < 4; x++)
        for (unsigned t = 0; t < 4; t++)
            for (int mu = 0; mu < 4; mu++) {
                S.computeStaple(L, x, 1, 2, t, mu);
                const unsigned long site = x + 4 * (1 + 4 * (2 + 4 * t));
                const double dS = S.getDeltaAction(L[mu].m_sites[site], two);
                REQUIRE(std::fabs(dS + 36.0) <= 1e-15, "dS = %.17g", dS);
            }
    std::printf("PASS dS = -36 on the unit field\n");
    // the same through GLAC's Action interface (GpuAction over the stand-in base)
    Action *A = new glac_gpu::GpuAction<Action, SU3, Lattice<SU3>>(6.0, nullptr);
    SU3 &u = L[2].m_sites[5];
    A->computeStaple(L, 1, 1, 0, 0, 2);
    const double dS = A->getDeltaAction(u, two);
    delete A;
    REQUIRE(std::fabs(dS + 36.0) <= 1e-15, "dS through Action* = %.17g", dS);
    std::printf("PASS dS = -36 through GLAC's Action interface\n");
    return 0;
}

// GLAC's System loop on the device (glac_gpu::System with GLAC-interface
// correlators), against the oracle running the same chain: GLAC's random hot
// start, thermalisation and NCf x NCor sweeps on the same link streams, each
// configuration flowed NFlows steps and measured.
static int run_system(const std::string &dir)
{
    const std::vector<unsigned int> N = {4, 4, 4, 8};
    const int n[4] = {4, 4, 4, 8};
    const unsigned long V = 4 * 4 * 4 * 8;
    glac_gpu::RunParameters p;
    p.N = N;
    p.beta = 6.0;
    p.NTherm = 3;
    p.NCf = 3;
    p.NCor = 2;
    p.NFlows = 5;
    p.flowEpsilon = 0.02;
    p.metropolisSeed = 31;
    p.randomMatrixSeed = 41;
    p.configPrefix = dir + "/system";
    using Obs = glac_gpu::GpuObservables<MasterSamplerTopcXYZ, Lattice<SU3>>;
    glac_gpu::Device dev(p.N, p.precision, p.device);
    Obs chainObs(&dev, false, p.NCf, 8u), flowObs(&dev, true, p.NFlows + 1, 8u);
    Correlator *corr = &chainObs;  // GLAC's m_correlator is a Correlator*
    Obs *flowCorr = &flowObs;
    {
        glac_gpu::System<Lattice<SU3>, Correlator, Obs> system(&dev, p, corr, flowCorr);
        glac_gpu::System<Lattice<SU3>, Correlator, Obs> *sys = &system;
        sys->latticeSetup();
        sys->run();

        // the oracle's chain
        std::vector<double> U(4 * V * 18), W, qt(8);
        or_hot_start(U.data(), n, 41, OR_HOT_GLAC_RANDOM, 0.24);
        unsigned long long sweep = 0;
        for (; sweep < 3; sweep++) or_cb_sweep(U.data(), n, 6.0, 0.24, 10, 31, sweep, 1);
        double acc = 0.0, out[3];
        for (unsigned c = 0; c < p.NCf; c++) {
            for (unsigned k = 0; k < p.NCor; k++) acc += or_cb_sweep(U.data(), n, 6.0, 0.24, 10, 31, sweep++, 1);
            W = U;
            REQUIRE(flowCorr->m_flowFiles.size() > c, "no flow file for configuration %u", c);
            for (unsigned i = 0; i <= p.NFlows; i++) {
                if (i > 0) or_flow_step(W.data(), n, 0.02);
                or_measure(W.data(), n, out, qt.data());
                const std::vector<double> &o = flowCorr->m_flowFiles[c][i];
                REQUIRE(std::fabs(o[0] - out[0]) < 1e-12, "config %u flow %u plaq %g vs %g", c, i, o[0], out[0]);
                REQUIRE(std::fabs(o[1] - out[2]) < 1e-12, "config %u flow %u topc", c, i);
                REQUIRE(std::fabs(o[2] / out[1] - 1.0) < 1e-11, "config %u flow %u energy", c, i);
                for (int t = 0; t < 8; t++) REQUIRE(std::fabs(o[3 + t] - qt[t]) < 1e-12, "config %u topct", c);
            }
            const std::vector<double> o0 = corr->getObservablesVector(c);
            REQUIRE(o0 == flowCorr->m_flowFiles[c][0], "chain observable %u is not the flow's t = 0 row", c);
        }
        acc /= double(p.NCf * p.NCor);
        REQUIRE(std::fabs(sys->getAcceptanceRate() - acc) < 1e-12, "acceptance %.15g vs %.15g",
                sys->getAcceptanceRate(), acc);
        REQUIRE(std::fabs(corr->m_acceptanceWritten - acc) < 1e-12, "acceptance handed to writeObservableToFile");
        // the chain after the run, det = -1 links included, and its last configuration file
        sys->syncHost();
        double e = 0, d[2], dmax = 0;
        for (int mu = 0; mu < 4; mu++)
            for (unsigned long s = 0; s < V; s++) {
                for (int k = 0; k < 18; k++)
                    e = std::fmax(e, std::fabs(sys->lattice()[mu].m_sites[s].mat[k] - U[(mu * V + s) * 18 + k]));
                or_su3_det(sys->lattice()[mu].m_sites[s].mat, d);
                dmax = std::fmax(dmax, std::fabs(d[0] + 1.0) + std::fabs(d[1]));
            }
        REQUIRE(e < 1e-10, "chain after the run differs from the oracle by %g", e);
        REQUIRE(dmax < 1e-12, "det = -1 not kept: %g", dmax);
        FILE *fp = std::fopen((dir + "/system_config00002.bin").c_str(), "rb");
        REQUIRE(fp, "configuration file not written");
        std::vector<double> file(V * 72);
        const size_t got = std::fread(file.data(), sizeof(double), file.size(), fp);
        std::fclose(fp);
        REQUIRE(got == file.size(), "short configuration file");
        e = 0;
        for (unsigned long s = 0; s < V; s++)
            for (int mu = 0; mu < 4; mu++)
                for (int k = 0; k < 18; k++)
                    e = std::fmax(e, std::fabs(file[(s * 4 + mu) * 18 + k] - U[(mu * V + s) * 18 + k]));
        REQUIRE(e < 1e-10, "configuration file differs from the oracle chain by %g", e);
    }
    std::printf("PASS System run loop (GLAC random start, %u + %u x %u sweeps, %u flow steps per configuration)\n",
                p.NTherm, p.NCf, p.NCor, p.NFlows);
    return 0;
}

static int run_gpu(const std::string &dir)
{
    const std::vector<unsigned int> N = {4, 4, 4, 8};
    const int n[4] = {4, 4, 4, 8};
    const unsigned long V = 4 * 4 * 4 * 8;
    Lattice<SU3> L[4];
    for (int mu = 0; mu < 4; mu++) L[mu].allocate(N);
    // oracle Haar field in GLAC's layout
    std::vector<double> U(4 * V * 18);
    or_hot_start_haar(U.data(), n, 77);
    for (int mu = 0; mu < 4; mu++)
        for (unsigned long s = 0; s < V; s++)
            for (int k = 0; k < 18; k++) L[mu].m_sites[s].mat[k] = U[(mu * V + s) * 18 + k];

    glac_gpu::Device dev(N, GLACX_FP64, 0);
    glac_gpu::WilsonGaugeAction<SU3, Lattice<SU3>> S(6.0, &dev);
    glac_gpu::Flow<Lattice<SU3>> flow(&dev, 0.02);
    glac_gpu::MasterSamplerTopcXYZ<Lattice<SU3>> obs(&dev, 6);

    // getActionDerivative vs the oracle
    for (int mu = 0; mu < 4; mu++) {
        Lattice<SU3> Z = S.getActionDerivative(L, mu);
        std::vector<double> Zo(V * 18);
        or_action_derivative(U.data(), n, mu, Zo.data());
        double err = 0;
        for (unsigned long s = 0; s < V; s++)
            for (int k = 0; k < 18; k++) err = std::fmax(err, std::fabs(Z.m_sites[s].mat[k] - Zo[s * 18 + k]));
        REQUIRE(err < 1e-13, "getActionDerivative mu %d err %g", mu, err);
    }
    std::printf("PASS getActionDerivative\n");

    // flowConfiguration loop: calculate(0); {flowField; calculate(i+1)} x 5
    obs.calculate(L, 0);
    for (int i = 0; i < 5; i++) {
        flow.flowField(L);
        obs.calculate(L, i + 1);
    }
    std::vector<double> W = U, qt(8);
    double out[3];
    for (int i = 0; i <= 5; i++) {
        if (i > 0) or_flow_step(W.data(), n, 0.02);
        or_measure(W.data(), n, out, qt.data());
        const std::vector<double> o = obs.getObservablesVector(i);
        REQUIRE(std::fabs(o[0] - out[0]) < 1e-12, "plaq step %d", i);
        REQUIRE(std::fabs(o[1] - out[2]) < 1e-12, "topc step %d", i);
        REQUIRE(std::fabs(o[2] / out[1] - 1.0) < 1e-11, "energy step %d", i);
        for (int t = 0; t < 8; t++) REQUIRE(std::fabs(o[3 + t] - qt[t]) < 1e-12, "topct step %d t %d", i, t);
    }
    // the host copy is stale until syncHost
    REQUIRE(std::fabs(L[0].m_sites[0].mat[0] - U[0]) == 0.0, "host copy changed without syncHost");
    dev.syncHost(L);
    double err = 0;
    for (int mu = 0; mu < 4; mu++)
        for (unsigned long s = 0; s < V; s++)
            for (int k = 0; k < 18; k++) err = std::fmax(err, std::fabs(L[mu].m_sites[s].mat[k] - W[(mu * V + s) * 18 + k]));
    REQUIRE(err < 1e-12, "flowed field after syncHost err %g", err);
    std::printf("PASS flow + clover observables through Flow / MasterSamplerTopcXYZ\n");

    // the same loop in one call, FP64 and FP32 links (the latter measuring
    // inside each step's first stage), on fresh copies of the start field
    for (int prec = 0; prec < 2; prec++) {
        Lattice<SU3> F[4];
        for (int mu = 0; mu < 4; mu++) {
            F[mu].allocate(N);
            for (unsigned long s = 0; s < V; s++)
                for (int k = 0; k < 18; k++) F[mu].m_sites[s].mat[k] = U[(mu * V + s) * 18 + k];
        }
        glac_gpu::Device devf(N, prec == 0 ? GLACX_FP64 : GLACX_FP32_LINKS, 0);
        glac_gpu::MasterSamplerTopcXYZ<Lattice<SU3>> obsf(&devf, 6);
        obsf.flowConfiguration(F, 5, 0.02);
        const double tp = prec == 0 ? 1e-12 : 5e-7, te = prec == 0 ? 1e-11 : 1e-5, tq = prec == 0 ? 1e-12 : 1e-4;
        std::vector<double> Wf = U;
        for (int i = 0; i <= 5; i++) {
            if (i > 0) or_flow_step(Wf.data(), n, 0.02);
            or_measure(Wf.data(), n, out, qt.data());
            const std::vector<double> o = obsf.getObservablesVector(i);
            REQUIRE(std::fabs(o[0] - out[0]) < tp, "flowConfiguration prec %d plaq step %d", prec, i);
            REQUIRE(std::fabs(o[1] - out[2]) < tq, "flowConfiguration prec %d topc step %d", prec, i);
            REQUIRE(std::fabs(o[2] / out[1] - 1.0) < te, "flowConfiguration prec %d energy step %d", prec, i);
            for (int t = 0; t < 8; t++)
                REQUIRE(std::fabs(o[3 + t] - qt[t]) < tq, "flowConfiguration prec %d topct step %d", prec, i);
        }
    }
    std::printf("PASS flowConfiguration in one call (FP64, FP32)\n");

    // Metropolis through System::update's interface on a separate chain
    Lattice<SU3> M[4];
    for (int mu = 0; mu < 4; mu++) M[mu].allocate(N);
    set_unit(M);
    glac_gpu::Device dev2(N, GLACX_FP64, 0);
    glac_gpu::Metropolis<Lattice<SU3>> metro(&dev2, 6.0, 0.24, 10, 2024);
    for (int s = 0; s < 20; s++) metro.update(M);
    dev2.syncHost(M);
    const double acc = metro.acceptanceScore() / 20.0;
    REQUIRE(acc > 0.05 && acc < 0.5, "acceptance %g", acc);
    double dmax = 0;
    for (int mu = 0; mu < 4; mu++)
        for (auto &s : M[mu].m_sites) {
            double d[2];
            or_su3_det(s.mat, d);
            dmax = std::fmax(dmax, std::fabs(d[0] - 1.0) + std::fabs(d[1]));
        }
    REQUIRE(dmax < 1e-12, "det deviation %g", dmax);
    std::printf("PASS Metropolis update, acceptance %.4f\n", acc);

    // FieldIO: write the chain through the device, read it back into another
    // host lattice, then the same data as a big-endian Chroma file
    const std::string bin = dir + "/adapter_cfg.bin", chroma = dir + "/adapter_cfg.chroma";
    glac_gpu::FieldIO<Lattice<SU3>> fio(&dev2);
    fio.writeFieldToFile(bin, M);
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1) {
            FILE *in = std::fopen(bin.c_str(), "rb");
            FILE *out = std::fopen(chroma.c_str(), "wb");
            REQUIRE(in && out, "cannot open %s / %s", bin.c_str(), chroma.c_str());
            unsigned char b[8], r[8];
            while (std::fread(b, 1, 8, in) == 8) {
                for (int k = 0; k < 8; k++) r[k] = b[7 - k];
                std::fwrite(r, 1, 8, out);
            }
            std::fclose(in);
            std::fclose(out);
        }
        Lattice<SU3> R[4];
        for (int mu = 0; mu < 4; mu++) R[mu].allocate(N);
        if (pass == 0)
            fio.loadFieldConfiguration(bin, R);
        else
            fio.loadChromaFieldConfiguration(chroma, R);
        dev2.syncHost(R);
        double e = 0;
        for (int mu = 0; mu < 4; mu++)
            for (unsigned long s = 0; s < V; s++)
                for (int k = 0; k < 18; k++) e = std::fmax(e, std::fabs(R[mu].m_sites[s].mat[k] - M[mu].m_sites[s].mat[k]));
        REQUIRE(e == 0.0, "%s round trip differs by %g", pass ? "chroma" : "bin", e);
    }
    std::printf("PASS FieldIO .bin and Chroma round trips\n");
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    if (mode == "cpu") return run_cpu();
    if (mode == "gpu") {
        const std::string dir = argc > 2 ? argv[2] : "/tmp";
        const int rc = run_gpu(dir);
        return rc ? rc : run_system(dir);
    }
    return 2;
}
