// api.hip -- the C ABI of libglacx.so (declared in include/glacx.h): context,
// device memory, host transfers, .bin I/O and the sweep / flow / measurement
// drivers that sequence the kernels on the context's stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/glacx.h"
#include "kernels.hpp"

using namespace glacx;

struct glacx_ctx {
    int dims[4] = {0, 0, 0, 0};
    Geom geom{};
    int precision = GLACX_FP64;
    int device = 0;
    size_t real_size = 8;
    hipStream_t stream = nullptr;
    void *field[2] = {nullptr, nullptr};
    void *work = nullptr;            // flow ping-pong buffer (one field)
    void *acc = nullptr;             // flow accumulator, 32 reals per site
    double *partials = nullptr;      // reduction partials
    size_t partials_len = 0;
    double *results = nullptr;       // device results (measurement rows)
    size_t results_len = 0;
    unsigned long long *counts = nullptr;  // accepted hits per sweep
    size_t counts_len = 0;
    double *stage = nullptr;         // AoS FP64 staging for host transfers
    double *scratch = nullptr;       // this device's copy of `scratch` (allocation-free calls)
    size_t device_bytes = 0;
    std::string err;
    // per-kernel event timing
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> timed;
    long long t_launches[GLACX_K_COUNT] = {};
    double t_ms[GLACX_K_COUNT] = {};
    // learned proposal
    float *samples = nullptr;        // column-major (A, U) pairs
    size_t samples_cap = 0, samples_n = 0;
    float *params = nullptr;         // FP32 master parameters, then Adam m, v, gradient
    float *adam_m = nullptr, *adam_v = nullptr, *grad = nullptr;
    unsigned short *params_bf16 = nullptr;  // inference copy
    float *mean_std = nullptr;       // 12 floats
    bool proposal_ready = false;
    long long adam_t = 0;
    float *inter = nullptr;          // per-sample training columns
    size_t inter_len = 0;
};

namespace {

int fail(glacx_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIP_TRY(ctx, call)                                                                              \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            return fail((ctx), GLACX_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));          \
    } while (0)

#define CHECK_LAUNCH(ctx) HIP_TRY(ctx, hipGetLastError())

size_t field_bytes(const glacx_ctx *c) { return (size_t)36 * c->geom.v * 2 * c->real_size; }

bool valid_field(int f) { return f == GLACX_FIELD_MAIN || f == GLACX_FIELD_FLOW; }

int dev_alloc(glacx_ctx *c, void **p, size_t bytes)
{
    HIP_TRY(c, hipMalloc(p, bytes));
    c->device_bytes += bytes;
    return GLACX_OK;
}

// Grow a device buffer of doubles to at least n entries.
template <class P> int ensure(glacx_ctx *c, P **buf, size_t *len, size_t n)
{
    if (*len >= n) return GLACX_OK;
    if (*buf) {
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        HIP_TRY(c, hipFree(*buf));
        c->device_bytes -= *len * sizeof(P);
        *buf = nullptr;
        *len = 0;
    }
    void *p = nullptr;
    int rc = dev_alloc(c, &p, n * sizeof(P));
    if (rc) return rc;
    *buf = (P *)p;
    *len = n;
    return GLACX_OK;
}

int ensure_stage(glacx_ctx *c)
{
    if (c->stage) return GLACX_OK;
    void *p = nullptr;
    int rc = dev_alloc(c, &p, (size_t)4 * c->geom.v * 18 * sizeof(double));
    c->stage = (double *)p;
    return rc;
}

// ---- event timing --------------------------------------------------------
int timed_begin(glacx_ctx *c, hipEvent_t *start)
{
    if (!c->timing) return GLACX_OK;
    if (c->ev_used + 2 > c->ev_pool.size()) {
        for (int k = 0; k < 64; k++) {
            hipEvent_t e;
            HIP_TRY(c, hipEventCreate(&e));
            c->ev_pool.push_back(e);
        }
    }
    *start = c->ev_pool[c->ev_used++];
    HIP_TRY(c, hipEventRecord(*start, c->stream));
    return GLACX_OK;
}

int timed_end(glacx_ctx *c, int kernel, hipEvent_t start)
{
    if (!c->timing) return GLACX_OK;
    hipEvent_t stop = c->ev_pool[c->ev_used++];
    HIP_TRY(c, hipEventRecord(stop, c->stream));
    c->timed.push_back({kernel, {start, stop}});
    // keep the pool bounded: fold finished pairs into the totals
    if (c->timed.size() >= 4096) {
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        for (auto &t : c->timed) {
            float ms = 0.f;
            HIP_TRY(c, hipEventElapsedTime(&ms, t.second.first, t.second.second));
            c->t_launches[t.first]++;
            c->t_ms[t.first] += ms;
        }
        c->timed.clear();
        c->ev_used = 0;
    }
    return GLACX_OK;
}

// ---- kernel sequences ----------------------------------------------------
// mode: 0 plain clover observables, 1 with the Weinberg operator, 2 with
// per-site densities written to `density` (2V doubles).
// `partials` holds measure_partials_count(geom) doubles.
template <class T, int R>
int run_measure_into(glacx_ctx *c, const void *f, double *partials, double *out, int mode, double *density = nullptr)
{
    hipEvent_t e0 = nullptr;
    int rc;
    if ((rc = timed_begin(c, &e0))) return rc;
    launch_measure<T, R>(f, c->geom, partials, mode, density, c->stream);
    CHECK_LAUNCH(c);
    if ((rc = timed_end(c, GLACX_K_MEASURE, e0))) return rc;
    launch_measure_finish(partials, c->geom, out, mode == 1, c->stream);
    CHECK_LAUNCH(c);
    return GLACX_OK;
}

template <class T, int R> int run_measure(glacx_ctx *c, const void *f, double *out, int mode, double *density = nullptr)
{
    int rc = ensure(c, &c->partials, &c->partials_len, (size_t)measure_partials_count(c->geom));
    if (rc) return rc;
    return run_measure_into<T, R>(c, f, c->partials, out, mode, density);
}

template <class T, int R> int run_plaquette(glacx_ctx *c, const void *f, double *out)
{
    const int np = plaquette_partials_count(c->geom);
    int rc = ensure(c, &c->partials, &c->partials_len, (size_t)np);
    if (rc) return rc;
    hipEvent_t e0 = nullptr;
    if ((rc = timed_begin(c, &e0))) return rc;
    launch_plaquette<T, R>(f, c->geom, c->partials, c->stream);
    CHECK_LAUNCH(c);
    if ((rc = timed_end(c, GLACX_K_PLAQUETTE, e0))) return rc;
    launch_plaquette_finish(c->partials, np, c->geom, out, c->stream);
    CHECK_LAUNCH(c);
    return GLACX_OK;
}

// One RK3 step on `f`: three stages ping-ponging through the work buffer
// (the caller swaps the roles of the two afterwards).  With `meas` stage 0
// also measures the clover observables of `f` into the measurement partials.
template <class T, int R> int run_flow_step(glacx_ctx *c, void *f, double eps, double *meas)
{
    void *bufs[2] = {f, c->work};
    int rc;
    for (int s = 0; s < 3; s++) {
        hipEvent_t e0 = nullptr;
        if ((rc = timed_begin(c, &e0))) return rc;
        launch_flow_stage<T, R>(bufs[s & 1], bufs[(s + 1) & 1], c->acc, c->geom, s, eps, s == 0 ? meas : nullptr,
                                c->stream);
        CHECK_LAUNCH(c);
        if ((rc = timed_end(c, s == 0 && meas ? GLACX_K_FLOW_STAGE_MEASURE : GLACX_K_FLOW_STAGE, e0))) return rc;
    }
    return GLACX_OK;
}

// n_steps flow steps of `field`, measuring before the first and after every
// step into results (rows of `row` doubles) when results is set, with
// `partials` (measure_partials_count(geom) doubles) as the reduction buffer.
// Each step's stage 0 measures the field it reads (the previous step's
// result), and the clover kernel measures once more after the last step; the
// SuperSampler set and lattices whose runs are not whole 32-site chunks use
// the clover kernel between steps.
template <class T, int R>
int run_flow(glacx_ctx *c, int field, int n_steps, double eps, double *results, size_t row, bool super,
             double *partials)
{
    const bool fuse = results && !super && flow_measure_fusable(c->geom);
    const int mode = super ? 1 : 0;
    int rc;
    if (results && !fuse && (rc = run_measure_into<T, R>(c, c->field[field], partials, results, mode))) return rc;
    for (int i = 0; i < n_steps; i++) {
        if ((rc = run_flow_step<T, R>(c, c->field[field], eps, fuse ? partials : nullptr))) return rc;
        if (fuse) {
            launch_measure_finish(partials, c->geom, results + row * (size_t)i, false, c->stream);
            CHECK_LAUNCH(c);
        }
        // three stages leave the result in the work buffer: swap roles
        std::swap(c->field[field], c->work);
        if (results && !fuse &&
            (rc = run_measure_into<T, R>(c, c->field[field], partials, results + row * (size_t)(i + 1), mode)))
            return rc;
    }
    if (fuse && (rc = run_measure_into<T, R>(c, c->field[field], partials, results + row * (size_t)n_steps, mode)))
        return rc;
    return GLACX_OK;
}

template <class T, int R> int run_metropolis_sweep(glacx_ctx *c, const MetroParams &p, unsigned long long *count)
{
    int rc;
    for (int mu = 0; mu < 4; mu++)
        for (int parity = 0; parity < 2; parity++) {
            hipEvent_t e0 = nullptr;
            if ((rc = timed_begin(c, &e0))) return rc;
            launch_metropolis<T, R>(c->field[GLACX_FIELD_MAIN], c->geom, mu, parity, p, count, c->stream);
            CHECK_LAUNCH(c);
            if ((rc = timed_end(c, GLACX_K_METROPOLIS, e0))) return rc;
        }
    return GLACX_OK;
}

int ensure_proposal_buffers(glacx_ctx *c)
{
    if (c->params) return GLACX_OK;
    void *p = nullptr;
    int rc = dev_alloc(c, &p, sizeof(float) * 4 * kProposalParams);  // params, m, v, grad
    if (rc) return rc;
    c->params = (float *)p;
    c->adam_m = c->params + kProposalParams;
    c->adam_v = c->adam_m + kProposalParams;
    c->grad = c->adam_v + kProposalParams;
    HIP_TRY(c, hipMemsetAsync(c->params, 0, sizeof(float) * 4 * kProposalParams, c->stream));
    if ((rc = dev_alloc(c, &p, sizeof(unsigned short) * kProposalParams))) return rc;
    c->params_bf16 = (unsigned short *)p;
    if ((rc = dev_alloc(c, &p, sizeof(float) * 12))) return rc;
    c->mean_std = (float *)p;
    return GLACX_OK;
}

template <class T, int R> int run_metropolis_learned_sweep(glacx_ctx *c, const MetroParams &p, unsigned long long *count)
{
    int rc;
    for (int mu = 0; mu < 4; mu++)
        for (int parity = 0; parity < 2; parity++) {
            hipEvent_t e0 = nullptr;
            if ((rc = timed_begin(c, &e0))) return rc;
            launch_metropolis_learned<T, R>(c->field[GLACX_FIELD_MAIN], c->geom, mu, parity, p, c->params_bf16,
                                         c->mean_std, count, c->stream);
            CHECK_LAUNCH(c);
            if ((rc = timed_end(c, GLACX_K_METROPOLIS_LEARNED, e0))) return rc;
        }
    return GLACX_OK;
}

template <class T, int R> int run_collect(glacx_ctx *c, int stride)
{
    const size_t per = (size_t)((c->geom.vh + stride - 1) / stride);
    for (int mu = 0; mu < 4; mu++)
        for (int parity = 0; parity < 2; parity++) {
            if (c->samples_n >= c->samples_cap) return GLACX_OK;
            launch_collect<T, R>(c->field[GLACX_FIELD_MAIN], c->geom, mu, parity, stride, c->samples, c->samples_cap,
                              c->samples_n, c->stream);
            CHECK_LAUNCH(c);
            c->samples_n = std::min(c->samples_cap, c->samples_n + per);
        }
    return GLACX_OK;
}

}  // namespace

#define CTX_CHECK(c)                                                                                    \
    do {                                                                                                 \
        if (!(c)) return GLACX_E_ARG;                                                                    \
        HIP_TRY(c, hipSetDevice((c)->device));                                                           \
    } while (0)

// Kernel instantiation for the context's precision: <real type, stored rows>.
#define DISPATCH(c, fn, ...)                                                                             \
    ((c)->precision == GLACX_FP32_LINKS        ? fn<float, 3>(__VA_ARGS__)                                \
     : (c)->precision == GLACX_FP32_COMPRESSED ? fn<float, 2>(__VA_ARGS__)                                \
                                               : fn<double, 3>(__VA_ARGS__))

namespace {

// The compressed mode stores two rows and rebuilds the third, which is only
// right for SU(3) links; refuse anything else (e.g. det = -1 links) at the
// door.  `link` points at 18 doubles in GLAC's layout.
double su3_defect(const double *m)
{
    double worst = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double re = 0.0, im = 0.0;
            for (int k = 0; k < 3; k++) {
                re += m[6 * i + 2 * k] * m[6 * j + 2 * k] + m[6 * i + 2 * k + 1] * m[6 * j + 2 * k + 1];
                im += m[6 * i + 2 * k + 1] * m[6 * j + 2 * k] - m[6 * i + 2 * k] * m[6 * j + 2 * k + 1];
            }
            worst = std::max(worst, std::fabs(re - (i == j ? 1.0 : 0.0)) + std::fabs(im));
        }
    // third row against conj(row0 x row1)
    for (int j = 0; j < 3; j++) {
        const int k = (j + 1) % 3, l = (j + 2) % 3;
        const double ar = m[2 * k], ai = m[2 * k + 1], br = m[6 + 2 * l], bi = m[6 + 2 * l + 1];
        const double cr = m[2 * l], ci = m[2 * l + 1], dr = m[6 + 2 * k], di = m[6 + 2 * k + 1];
        const double re = ar * br - ai * bi - (cr * dr - ci * di);
        const double im = ar * bi + ai * br - (cr * di + ci * dr);
        worst = std::max(worst, std::fabs(m[12 + 2 * j] - re) + std::fabs(m[12 + 2 * j + 1] + im));
    }
    return worst;
}

int require_su3(glacx_ctx *c, const double *links, size_t n_links, size_t stride, const char *who)
{
    if (c->precision != GLACX_FP32_COMPRESSED) return GLACX_OK;
    for (size_t i = 0; i < n_links; i++) {
        const double d = su3_defect(links + i * stride);
        if (!(d < 1e-6))
            return fail(c, GLACX_E_ARG, std::string(who) + ": the compressed FP32 mode stores two rows per link and "
                                        "needs SU(3) links; link " + std::to_string(i) + " deviates by " +
                                        std::to_string(d));
    }
    return GLACX_OK;
}

}  // namespace

// ---- scratch memory of the allocation-free entry point -------------------
// glacx_flow_noalloc keeps its reduction partials and per-step results in
// this fixed array instead of context buffers that grow on demand, so the
// call makes no device allocation.  Every GPU holds its own copy for the
// life of the process; host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (the array's name is a
// host-side stand-in, not a device address).
constexpr size_t kScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch_addr[kMaxDevices] = {};

// Device address of `scratch` on c->device (current), looked up once per
// device; called from glacx_create.
int scratch_setup(glacx_ctx *c)
{
    if (c->device < 0 || c->device >= kMaxDevices)
        return fail(c, GLACX_E_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_addr[c->device]) {
        void *p = nullptr;
        HIP_TRY(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(c, GLACX_E_HIP, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch_addr[c->device] = (double *)p;
    }
    c->scratch = g_scratch_addr[c->device];
    return GLACX_OK;
}

}  // namespace

extern "C" {

int glacx_create(const int dims[4], int precision, int device, glacx_ctx **out)
{
    if (!dims || !out) return GLACX_E_ARG;
    *out = nullptr;
    long v = 1;
    for (int k = 0; k < 4; k++) {
        if (dims[k] < 2 || (dims[k] & 1) || dims[k] > 512) return GLACX_E_ARG;
        v *= dims[k];
    }
    // the 9 planes of one direction are addressed by one buffer resource
    if (precision != GLACX_FP64 && precision != GLACX_FP32_LINKS && precision != GLACX_FP32_COMPRESSED)
        return GLACX_E_ARG;
    if ((unsigned long long)v * 9ull * 2ull * (precision != GLACX_FP64 ? 4ull : 8ull) >= (1ull << 32))
        return GLACX_E_ARG;
    glacx_ctx *c = new glacx_ctx;
    for (int k = 0; k < 4; k++) c->dims[k] = dims[k];
    c->geom = make_geom(dims);
    c->precision = precision;
    c->real_size = precision != GLACX_FP64 ? 4 : 8;
    c->device = device;
    int rc = GLACX_OK;
    auto bad = [&](int code) {
        *out = c;  // keep the context so the caller can read the message
        return code;
    };
    if (hipSetDevice(device) != hipSuccess) return bad(fail(c, GLACX_E_HIP, "hipSetDevice failed"));
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bad(fail(c, GLACX_E_HIP, "hipStreamCreate failed"));
    for (int f = 0; f < 2; f++)
        if ((rc = dev_alloc(c, &c->field[f], field_bytes(c)))) return bad(rc);
    if ((rc = dev_alloc(c, &c->work, field_bytes(c)))) return bad(rc);
    if ((rc = dev_alloc(c, &c->acc, (size_t)32 * v * c->real_size))) return bad(rc);
    if ((rc = ensure(c, &c->results, &c->results_len, (size_t)(3 + dims[3])))) return bad(rc);
    if ((rc = scratch_setup(c))) return bad(rc);
    *out = c;
    return GLACX_OK;
}

int glacx_destroy(glacx_ctx *c)
{
    if (!c) return GLACX_E_ARG;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void *p : {c->field[0], c->field[1], c->work, c->acc, (void *)c->partials, (void *)c->results,
                    (void *)c->counts, (void *)c->stage, (void *)c->samples, (void *)c->params,
                    (void *)c->params_bf16, (void *)c->mean_std, (void *)c->inter})
        if (p) (void)hipFree(p);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return GLACX_OK;
}

const char *glacx_last_error(const glacx_ctx *c) { return c ? c->err.c_str() : "null context"; }

int glacx_device_bytes(const glacx_ctx *c, size_t *bytes)
{
    if (!c || !bytes) return GLACX_E_ARG;
    *bytes = c->device_bytes;
    return GLACX_OK;
}

int glacx_synchronize(glacx_ctx *c)
{
    CTX_CHECK(c);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return GLACX_OK;
}

int glacx_upload_field(glacx_ctx *c, int field, const double *const links[4])
{
    CTX_CHECK(c);
    if (!valid_field(field) || !links) return fail(c, GLACX_E_ARG, "upload: bad field or pointer");
    for (int mu = 0; mu < 4; mu++)
        if (!links[mu]) return fail(c, GLACX_E_ARG, "upload: null link array");
    int rc;
    for (int mu = 0; mu < 4; mu++)
        if ((rc = require_su3(c, links[mu], (size_t)c->geom.v, 18, "upload"))) return rc;
    if ((rc = ensure_stage(c))) return rc;
    const size_t n = (size_t)c->geom.v * 18;
    for (int mu = 0; mu < 4; mu++)
        HIP_TRY(c, hipMemcpyAsync(c->stage + mu * n, links[mu], n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    DISPATCH(c, launch_pack, c->stage, c->field[field], c->geom, 0, c->stream);
    CHECK_LAUNCH(c);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return GLACX_OK;
}

int glacx_download_field(glacx_ctx *c, int field, double *const links[4])
{
    CTX_CHECK(c);
    if (!valid_field(field) || !links) return fail(c, GLACX_E_ARG, "download: bad field or pointer");
    for (int mu = 0; mu < 4; mu++)
        if (!links[mu]) return fail(c, GLACX_E_ARG, "download: null link array");
    int rc = ensure_stage(c);
    if (rc) return rc;
    DISPATCH(c, launch_unpack, c->field[field], c->stage, c->geom, 0, c->stream);
    CHECK_LAUNCH(c);
    const size_t n = (size_t)c->geom.v * 18;
    for (int mu = 0; mu < 4; mu++)
        HIP_TRY(c, hipMemcpyAsync(links[mu], c->stage + mu * n, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return GLACX_OK;
}

int glacx_copy_field(glacx_ctx *c, int dst, int src)
{
    CTX_CHECK(c);
    if (!valid_field(dst) || !valid_field(src)) return fail(c, GLACX_E_ARG, "copy: bad field");
    if (dst == src) return GLACX_OK;
    HIP_TRY(c, hipMemcpyAsync(c->field[dst], c->field[src], field_bytes(c), hipMemcpyDeviceToDevice, c->stream));
    return GLACX_OK;
}

namespace {

// Reads V * 72 doubles of GLAC's .bin site order and uploads them; with
// `byteswap` every double is converted from big-endian first (Chroma files).
int load_site_major(glacx_ctx *c, int field, const char *path, bool byteswap, const char *who)
{
    if (!valid_field(field) || !path) return fail(c, GLACX_E_ARG, std::string(who) + ": bad field or path");
    const size_t n = (size_t)c->geom.v * 72;
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return fail(c, GLACX_E_IO, std::string(who) + ": cannot open " + path);
    std::vector<double> buf(n);
    const size_t got = std::fread(buf.data(), sizeof(double), n, fp);
    std::fclose(fp);
    if (got != n) return fail(c, GLACX_E_IO, std::string(who) + ": short file " + path);
    if (byteswap)
        for (size_t i = 0; i < n; i++) {
            uint64_t u;
            std::memcpy(&u, &buf[i], sizeof u);
            u = __builtin_bswap64(u);
            std::memcpy(&buf[i], &u, sizeof u);
        }
    int rc = require_su3(c, buf.data(), (size_t)c->geom.v * 4, 18, who);
    if (rc) return rc;
    if ((rc = ensure_stage(c))) return rc;
    HIP_TRY(c, hipMemcpyAsync(c->stage, buf.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    DISPATCH(c, launch_pack, c->stage, c->field[field], c->geom, 1, c->stream);
    CHECK_LAUNCH(c);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return GLACX_OK;
}

}  // namespace

int glacx_load_bin(glacx_ctx *c, int field, const char *path)
{
    CTX_CHECK(c);
    return load_site_major(c, field, path, false, "load_bin");
}

int glacx_load_chroma(glacx_ctx *c, int field, const char *path)
{
    CTX_CHECK(c);
    return load_site_major(c, field, path, true, "load_chroma");
}

int glacx_save_bin(glacx_ctx *c, int field, const char *path)
{
    CTX_CHECK(c);
    if (!valid_field(field) || !path) return fail(c, GLACX_E_ARG, "save_bin: bad field or path");
    int rc = ensure_stage(c);
    if (rc) return rc;
    const size_t n = (size_t)c->geom.v * 72;
    std::vector<double> buf(n);
    DISPATCH(c, launch_unpack, c->field[field], c->stage, c->geom, 1, c->stream);
    CHECK_LAUNCH(c);
    HIP_TRY(c, hipMemcpyAsync(buf.data(), c->stage, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return fail(c, GLACX_E_IO, std::string("save_bin: cannot open ") + path);
    const size_t put = std::fwrite(buf.data(), sizeof(double), n, fp);
    const int closed = std::fclose(fp);
    if (put != n || closed != 0) return fail(c, GLACX_E_IO, std::string("save_bin: write failed ") + path);
    return GLACX_OK;
}

int glacx_cold_start(glacx_ctx *c, int field)
{
    CTX_CHECK(c);
    if (!valid_field(field)) return fail(c, GLACX_E_ARG, "cold_start: bad field");
    DISPATCH(c, launch_cold_start, c->field[field], c->geom, c->stream);
    CHECK_LAUNCH(c);
    return GLACX_OK;
}

int glacx_hot_start(glacx_ctx *c, int field, uint64_t seed, int kind, double eps)
{
    CTX_CHECK(c);
    if (!valid_field(field)) return fail(c, GLACX_E_ARG, "hot_start: bad field");
    if (kind != GLACX_HOT_HAAR && kind != GLACX_HOT_GLAC_RANDOM && kind != GLACX_HOT_RST)
        return fail(c, GLACX_E_ARG, "hot_start: unknown kind");
    if (kind != GLACX_HOT_HAAR && !(eps > 0.0 && eps < 1.0))
        return fail(c, GLACX_E_ARG, "hot_start: need 0 < eps < 1");
    if (kind == GLACX_HOT_GLAC_RANDOM && c->precision == GLACX_FP32_COMPRESSED)
        return fail(c, GLACX_E_ARG, "hot_start: GLAC's random start gives det U = -1 links, which the compressed "
                                    "FP32 mode (two stored rows, det = +1) cannot hold; use GLACX_FP32_LINKS");
    DISPATCH(c, launch_hot_start, c->field[field], c->geom, seed, kind, eps, c->stream);
    CHECK_LAUNCH(c);
    return GLACX_OK;
}

int glacx_metropolis(glacx_ctx *c, int n_sweeps, double beta, double eps, int n_hit, int reunit_every, uint64_t seed,
                     uint64_t sweep_offset, double *accept, double *plaq)
{
    CTX_CHECK(c);
    if (n_sweeps < 0 || n_hit < 1 || !(eps > 0.0 && eps < 1.0))
        return fail(c, GLACX_E_ARG, "metropolis: need n_sweeps >= 0, n_hit >= 1, 0 < eps < 1");
    if (n_sweeps == 0) return GLACX_OK;
    int rc = ensure(c, &c->counts, &c->counts_len, (size_t)n_sweeps);
    if (rc) return rc;
    HIP_TRY(c, hipMemsetAsync(c->counts, 0, sizeof(unsigned long long) * n_sweeps, c->stream));
    if (plaq && (rc = ensure(c, &c->results, &c->results_len, (size_t)n_sweeps))) return rc;
    for (int i = 0; i < n_sweeps; i++) {
        MetroParams p;
        p.beta = beta;
        p.eps = eps;
        p.n_hit = n_hit;
        p.sweep = sweep_offset + (uint64_t)i;
        p.reunit = reunit_every > 0 && (p.sweep % (uint64_t)reunit_every) == 0;
        p.seed = seed;
        if ((rc = DISPATCH(c, run_metropolis_sweep, c, p, c->counts + i))) return rc;
        if (plaq && (rc = DISPATCH(c, run_plaquette, c, c->field[GLACX_FIELD_MAIN], c->results + i))) return rc;
    }
    std::vector<unsigned long long> counts(n_sweeps);
    HIP_TRY(c, hipMemcpyAsync(counts.data(), c->counts, sizeof(unsigned long long) * n_sweeps, hipMemcpyDeviceToHost,
                              c->stream));
    if (plaq) HIP_TRY(c, hipMemcpyAsync(plaq, c->results, sizeof(double) * n_sweeps, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (accept)
        for (int i = 0; i < n_sweeps; i++) accept[i] = (double)counts[i] / (4.0 * n_hit * (double)c->geom.v);
    return GLACX_OK;
}

namespace {

// Clover observables of `field`; with `super` also GLAC's Weinberg operator.
int measure_impl(glacx_ctx *c, int field, bool super, double *plaq, double *e_raw, double *q, double *w, double *q_t,
                 double *w_t)
{
    if (!valid_field(field)) return fail(c, GLACX_E_ARG, "measure: bad field");
    const int nt = c->dims[3];
    const size_t row = super ? (size_t)4 + 2 * nt : (size_t)3 + nt;
    int rc = ensure(c, &c->results, &c->results_len, row);
    if (rc) return rc;
    if ((rc = DISPATCH(c, run_measure, c, c->field[field], c->results, super ? 1 : 0))) return rc;
    std::vector<double> h(row);
    HIP_TRY(c, hipMemcpyAsync(h.data(), c->results, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const int head = super ? 4 : 3;
    if (plaq) *plaq = h[0];
    if (e_raw) *e_raw = h[1];
    if (q) *q = h[2];
    if (w) *w = h[3];
    if (q_t) std::memcpy(q_t, h.data() + head, sizeof(double) * nt);
    if (w_t) std::memcpy(w_t, h.data() + head + nt, sizeof(double) * nt);
    return GLACX_OK;
}

}  // namespace

int glacx_measure(glacx_ctx *c, int field, double *plaq, double *e_raw, double *q, double *q_t)
{
    CTX_CHECK(c);
    return measure_impl(c, field, false, plaq, e_raw, q, nullptr, q_t, nullptr);
}

int glacx_measure_super(glacx_ctx *c, int field, double *plaq, double *e_raw, double *q, double *w, double *q_t,
                        double *w_t)
{
    CTX_CHECK(c);
    return measure_impl(c, field, true, plaq, e_raw, q, w, q_t, w_t);
}

int glacx_measure_density(glacx_ctx *c, int field, double *q_density, double *e_density, double *plaq,
                          double *e_raw, double *q)
{
    CTX_CHECK(c);
    if (!valid_field(field)) return fail(c, GLACX_E_ARG, "measure_density: bad field");
    int rc = ensure_stage(c);
    if (rc) return rc;
    if ((rc = ensure(c, &c->results, &c->results_len, (size_t)3 + c->dims[3]))) return rc;
    if ((rc = DISPATCH(c, run_measure, c, c->field[field], c->results, 2, c->stage))) return rc;
    const size_t v = (size_t)c->geom.v;
    double head[3];
    HIP_TRY(c, hipMemcpyAsync(head, c->results, sizeof head, hipMemcpyDeviceToHost, c->stream));
    if (q_density) HIP_TRY(c, hipMemcpyAsync(q_density, c->stage, v * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (e_density)
        HIP_TRY(c, hipMemcpyAsync(e_density, c->stage + v, v * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (plaq) *plaq = head[0];
    if (e_raw) *e_raw = head[1];
    if (q) *q = head[2];
    return GLACX_OK;
}

int glacx_plaquette(glacx_ctx *c, int field, double *plaq)
{
    CTX_CHECK(c);
    if (!valid_field(field) || !plaq) return fail(c, GLACX_E_ARG, "plaquette: bad field or pointer");
    int rc = DISPATCH(c, run_plaquette, c, c->field[field], c->results);
    if (rc) return rc;
    HIP_TRY(c, hipMemcpyAsync(plaq, c->results, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return GLACX_OK;
}

namespace {

int flow_impl(glacx_ctx *c, int field, int n_steps, double eps, bool super, double *plaq, double *e_raw, double *q,
              double *w, double *q_t, double *w_t)
{
    if (!valid_field(field) || n_steps < 0) return fail(c, GLACX_E_ARG, "flow: bad field or step count");
    const bool meas = plaq || e_raw || q || q_t || w || w_t;
    const int nt = c->dims[3];
    const size_t row = super ? (size_t)4 + 2 * nt : (size_t)3 + nt;
    int rc;
    if (meas && (rc = ensure(c, &c->results, &c->results_len, row * (size_t)(n_steps + 1)))) return rc;
    if (meas && (rc = ensure(c, &c->partials, &c->partials_len, (size_t)measure_partials_count(c->geom)))) return rc;
    if ((rc = DISPATCH(c, run_flow, c, field, n_steps, eps, meas ? c->results : nullptr, row, super, c->partials)))
        return rc;
    if (meas) {
        std::vector<double> h(row * (size_t)(n_steps + 1));
        HIP_TRY(c, hipMemcpyAsync(h.data(), c->results, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        const int head = super ? 4 : 3;
        for (int i = 0; i <= n_steps; i++) {
            const double *r = h.data() + row * i;
            if (plaq) plaq[i] = r[0];
            if (e_raw) e_raw[i] = r[1];
            if (q) q[i] = r[2];
            if (w) w[i] = r[3];
            if (q_t) std::memcpy(q_t + (size_t)i * nt, r + head, sizeof(double) * nt);
            if (w_t) std::memcpy(w_t + (size_t)i * nt, r + head + nt, sizeof(double) * nt);
        }
    } else {
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    return GLACX_OK;
}

}  // namespace

int glacx_flow(glacx_ctx *c, int field, int n_steps, double eps, double *plaq, double *e_raw, double *q, double *q_t)
{
    CTX_CHECK(c);
    return flow_impl(c, field, n_steps, eps, false, plaq, e_raw, q, nullptr, q_t, nullptr);
}

int glacx_flow_super(glacx_ctx *c, int field, int n_steps, double eps, double *plaq, double *e_raw, double *q,
                     double *w, double *q_t, double *w_t)
{
    CTX_CHECK(c);
    return flow_impl(c, field, n_steps, eps, true, plaq, e_raw, q, w, q_t, w_t);
}

int glacx_flow_noalloc(glacx_ctx *c, int field, int n_steps, double eps, double *plaq, double *e_raw, double *q,
                       double *q_t)
{
    CTX_CHECK(c);
    if (!valid_field(field) || n_steps < 0) return fail(c, GLACX_E_ARG, "flow_noalloc: bad field or step count");
    if (!c->scratch) return fail(c, GLACX_E_STATE, "flow_noalloc: no scratch address for this device");
    const bool meas = plaq || e_raw || q || q_t;
    const int nt = c->dims[3];
    const size_t row = (size_t)3 + nt;
    const size_t np = (size_t)measure_partials_count(c->geom);
    const size_t need = meas ? np + row * ((size_t)n_steps + 1) : 0;
    if (need * sizeof(double) > kScratchBytes)
        return fail(c, GLACX_E_ARG, "flow_noalloc: " + std::to_string(n_steps) + " measured steps on this lattice need " +
                                        std::to_string(need * sizeof(double)) + " bytes of scratch, above its " +
                                        std::to_string(kScratchBytes) + "; split the flow into shorter calls");
    double *partials = c->scratch, *results = c->scratch + np;
    // event timing would create events: off for this call
    const bool timing = c->timing;
    c->timing = false;
    int rc = DISPATCH(c, run_flow, c, field, n_steps, eps, meas ? results : nullptr, row, false, partials);
    c->timing = timing;
    if (rc) return rc;
    // rows {plaq, E_raw, Q, Q(t)...} to the caller's arrays by strided copies
    const size_t pitch = row * sizeof(double), rows = (size_t)n_steps + 1;
    double *cols[3] = {plaq, e_raw, q};
    for (int k = 0; k < 3; k++)
        if (cols[k])
            HIP_TRY(c, hipMemcpy2DAsync(cols[k], sizeof(double), results + k, pitch, sizeof(double), rows,
                                        hipMemcpyDeviceToHost, c->stream));
    if (q_t)
        HIP_TRY(c, hipMemcpy2DAsync(q_t, nt * sizeof(double), results + 3, pitch, nt * sizeof(double), rows,
                                    hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return GLACX_OK;
}

int glacx_action_derivative(glacx_ctx *c, int field, int mu, double *z_out)
{
    CTX_CHECK(c);
    if (!valid_field(field) || mu < 0 || mu > 3 || !z_out)
        return fail(c, GLACX_E_ARG, "action_derivative: bad field, mu or pointer");
    int rc = ensure_stage(c);
    if (rc) return rc;
    DISPATCH(c, launch_action_derivative, c->field[field], c->stage, c->geom, mu, c->stream);
    CHECK_LAUNCH(c);
    HIP_TRY(c, hipMemcpyAsync(z_out, c->stage, (size_t)c->geom.v * 18 * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return GLACX_OK;
}

int glacx_samples_reset(glacx_ctx *c, size_t capacity)
{
    CTX_CHECK(c);
    if (capacity == 0) return fail(c, GLACX_E_ARG, "samples_reset: zero capacity");
    if (c->samples) {
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        HIP_TRY(c, hipFree(c->samples));
        c->device_bytes -= sizeof(float) * 36 * c->samples_cap;
        c->samples = nullptr;
    }
    void *p = nullptr;
    int rc = dev_alloc(c, &p, sizeof(float) * 36 * capacity);
    if (rc) return rc;
    c->samples = (float *)p;
    c->samples_cap = capacity;
    c->samples_n = 0;
    return GLACX_OK;
}

int glacx_samples_collect(glacx_ctx *c, int stride, size_t *count)
{
    CTX_CHECK(c);
    if (!c->samples || stride < 1) return fail(c, GLACX_E_STATE, "samples_collect: no sample buffer or bad stride");
    int rc = DISPATCH(c, run_collect, c, stride);
    if (rc) return rc;
    if (count) *count = c->samples_n;
    return GLACX_OK;
}

int glacx_samples_stats(glacx_ctx *c, double beta, float mean[6], float stdv[6])
{
    CTX_CHECK(c);
    if (!c->samples || c->samples_n == 0) return fail(c, GLACX_E_STATE, "samples_stats: no samples");
    int rc = ensure_proposal_buffers(c);
    if (rc) return rc;
    if ((rc = ensure(c, &c->partials, &c->partials_len, (size_t)12 * feature_moments_blocks(c->samples_n)))) return rc;
    launch_feature_moments(c->samples, c->samples_cap, c->samples_n, beta, c->partials, c->mean_std, c->stream);
    CHECK_LAUNCH(c);
    float h[12];
    HIP_TRY(c, hipMemcpyAsync(h, c->mean_std, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (mean) std::memcpy(mean, h, 6 * sizeof(float));
    if (stdv) std::memcpy(stdv, h + 6, 6 * sizeof(float));
    return GLACX_OK;
}

int glacx_proposal_set(glacx_ctx *c, const float *params, const float mean[6], const float stdv[6])
{
    CTX_CHECK(c);
    if (!params || !mean || !stdv) return fail(c, GLACX_E_ARG, "proposal_set: null pointer");
    for (int k = 0; k < 6; k++)
        if (!(stdv[k] > 0.f)) return fail(c, GLACX_E_ARG, "proposal_set: std must be positive");
    int rc = ensure_proposal_buffers(c);
    if (rc) return rc;
    float ms[12];
    std::memcpy(ms, mean, 6 * sizeof(float));
    std::memcpy(ms + 6, stdv, 6 * sizeof(float));
    HIP_TRY(c, hipMemcpyAsync(c->params, params, sizeof(float) * kProposalParams, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->mean_std, ms, sizeof ms, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemsetAsync(c->adam_m, 0, sizeof(float) * 2 * kProposalParams, c->stream));
    launch_to_bf16(c->params, c->params_bf16, c->stream);
    CHECK_LAUNCH(c);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->adam_t = 0;
    c->proposal_ready = true;
    return GLACX_OK;
}

int glacx_proposal_get(glacx_ctx *c, float *params, float mean[6], float stdv[6])
{
    CTX_CHECK(c);
    if (!c->proposal_ready) return fail(c, GLACX_E_STATE, "proposal_get: no proposal set");
    float ms[12];
    if (params)
        HIP_TRY(c, hipMemcpyAsync(params, c->params, sizeof(float) * kProposalParams, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(ms, c->mean_std, sizeof ms, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (mean) std::memcpy(mean, ms, 6 * sizeof(float));
    if (stdv) std::memcpy(stdv, ms + 6, 6 * sizeof(float));
    return GLACX_OK;
}

int glacx_train(glacx_ctx *c, int steps, int batch, int draws, double lr, double beta, uint64_t seed, double *loss)
{
    CTX_CHECK(c);
    if (!c->proposal_ready) return fail(c, GLACX_E_STATE, "train: set initial parameters first");
    if (!c->samples || c->samples_n == 0) return fail(c, GLACX_E_STATE, "train: no samples");
    if (steps < 0 || batch < 1 || draws < 1 || !(lr > 0.0)) return fail(c, GLACX_E_ARG, "train: bad arguments");
    int rc;
    const int nblk = (batch + kBlock - 1) / kBlock;
    if ((rc = ensure(c, &c->inter, &c->inter_len, (size_t)kTrainInterCols * batch))) return rc;
    if ((rc = ensure(c, &c->partials, &c->partials_len, (size_t)nblk * steps))) return rc;
    for (int i = 0; i < steps; i++) {
        hipEvent_t e0 = nullptr;
        if ((rc = timed_begin(c, &e0))) return rc;
        c->adam_t++;
        launch_train_step(c->samples, c->samples_cap, c->samples_n, batch, draws, seed, (uint64_t)c->adam_t, beta,
                          c->params, c->mean_std, nullptr, c->inter, c->partials + (size_t)i * nblk, c->grad,
                          c->adam_m, c->adam_v, (float)lr, (int)c->adam_t, c->params_bf16, c->stream);
        CHECK_LAUNCH(c);
        if ((rc = timed_end(c, GLACX_K_TRAIN, e0))) return rc;
    }
    if (loss && steps > 0) {
        std::vector<double> h((size_t)nblk * steps);
        HIP_TRY(c, hipMemcpyAsync(h.data(), c->partials, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        for (int i = 0; i < steps; i++) {
            double s = 0.0;
            for (int b = 0; b < nblk; b++) s += h[(size_t)i * nblk + b];
            loss[i] = s;
        }
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return GLACX_OK;
}

int glacx_train_grad(glacx_ctx *c, const float *samples, int n, const float *z, int draws, double beta, double *loss,
                     float *grad)
{
    CTX_CHECK(c);
    if (!c->proposal_ready) return fail(c, GLACX_E_STATE, "train_grad: set parameters first");
    if (!samples || !z || n < 1 || draws < 1) return fail(c, GLACX_E_ARG, "train_grad: bad arguments");
    // column-major copy of the host samples, then the draws
    std::vector<float> cols((size_t)36 * n);
    for (int s = 0; s < n; s++)
        for (int k = 0; k < 36; k++) cols[(size_t)k * n + s] = samples[(size_t)s * 36 + k];
    float *dsmp = nullptr, *dz = nullptr;
    HIP_TRY(c, hipMalloc((void **)&dsmp, cols.size() * sizeof(float)));
    const size_t nz = (size_t)n * draws * 8;
    hipError_t e = hipMalloc((void **)&dz, nz * sizeof(float));
    if (e != hipSuccess) {
        (void)hipFree(dsmp);
        return fail(c, GLACX_E_HIP, "train_grad: hipMalloc failed");
    }
    int rc = GLACX_OK;
    const int nblk = (n + kBlock - 1) / kBlock;
    do {
        if ((rc = ensure(c, &c->inter, &c->inter_len, (size_t)kTrainInterCols * n))) break;
        if ((rc = ensure(c, &c->partials, &c->partials_len, (size_t)nblk))) break;
        if (hipMemcpyAsync(dsmp, cols.data(), cols.size() * sizeof(float), hipMemcpyHostToDevice, c->stream) !=
                hipSuccess ||
            hipMemcpyAsync(dz, z, nz * sizeof(float), hipMemcpyHostToDevice, c->stream) != hipSuccess) {
            rc = fail(c, GLACX_E_HIP, "train_grad: copy failed");
            break;
        }
        launch_train_step(dsmp, (size_t)n, (size_t)n, n, draws, 0, 0, beta, c->params, c->mean_std, dz, c->inter,
                          c->partials, c->grad, c->adam_m, c->adam_v, 0.f, 0, c->params_bf16, c->stream);
        if (hipGetLastError() != hipSuccess) {
            rc = fail(c, GLACX_E_HIP, "train_grad: launch failed");
            break;
        }
        std::vector<double> h(nblk);
        if (hipMemcpyAsync(h.data(), c->partials, nblk * sizeof(double), hipMemcpyDeviceToHost, c->stream) !=
                hipSuccess ||
            (grad && hipMemcpyAsync(grad, c->grad, sizeof(float) * kProposalParams, hipMemcpyDeviceToHost,
                                    c->stream) != hipSuccess) ||
            hipStreamSynchronize(c->stream) != hipSuccess) {
            rc = fail(c, GLACX_E_HIP, "train_grad: readback failed");
            break;
        }
        double s = 0.0;
        for (double v : h) s += v;
        if (loss) *loss = s;
    } while (0);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(dsmp);
    (void)hipFree(dz);
    return rc;
}

int glacx_proposal_eval(glacx_ctx *c, const float *features, int n, float *out, int *in_range)
{
    CTX_CHECK(c);
    if (!c->proposal_ready) return fail(c, GLACX_E_STATE, "proposal_eval: set parameters first");
    if (!features || !out || n < 1) return fail(c, GLACX_E_ARG, "proposal_eval: bad arguments");
    const size_t nf = (size_t)6 * n, no = (size_t)36 * n;
    void *buf = nullptr;
    HIP_TRY(c, hipMalloc(&buf, (nf + no) * sizeof(float) + (size_t)n * sizeof(int)));
    float *dfeat = (float *)buf, *dout = dfeat + nf;
    int *dok = (int *)(dout + no);
    int rc = GLACX_OK;
    if (hipMemcpyAsync(dfeat, features, nf * sizeof(float), hipMemcpyHostToDevice, c->stream) != hipSuccess) {
        rc = fail(c, GLACX_E_HIP, "proposal_eval: copy failed");
    } else {
        launch_proposal_eval(dfeat, n, c->params_bf16, c->mean_std, dout, dok, c->stream);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(out, dout, no * sizeof(float), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            (in_range && hipMemcpyAsync(in_range, dok, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, c->stream) !=
                             hipSuccess) ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            rc = fail(c, GLACX_E_HIP, "proposal_eval: launch or readback failed");
    }
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(buf);
    return rc;
}

int glacx_metropolis_learned(glacx_ctx *c, int n_sweeps, double beta, double eps, int n_hit, int reunit_every,
                             uint64_t seed, uint64_t sweep_offset, double *accept, double *plaq)
{
    CTX_CHECK(c);
    if (!c->proposal_ready) return fail(c, GLACX_E_STATE, "metropolis_learned: no proposal set");
    if (n_sweeps < 0 || n_hit < 1 || !(eps > 0.0 && eps < 1.0))
        return fail(c, GLACX_E_ARG, "metropolis_learned: need n_sweeps >= 0, n_hit >= 1, 0 < eps < 1");
    if (n_sweeps == 0) return GLACX_OK;
    int rc = ensure(c, &c->counts, &c->counts_len, (size_t)n_sweeps);
    if (rc) return rc;
    HIP_TRY(c, hipMemsetAsync(c->counts, 0, sizeof(unsigned long long) * n_sweeps, c->stream));
    if (plaq && (rc = ensure(c, &c->results, &c->results_len, (size_t)n_sweeps))) return rc;
    for (int i = 0; i < n_sweeps; i++) {
        MetroParams p;
        p.beta = beta;
        p.eps = eps;
        p.n_hit = n_hit;
        p.sweep = sweep_offset + (uint64_t)i;
        p.reunit = reunit_every > 0 && (p.sweep % (uint64_t)reunit_every) == 0;
        p.seed = seed;
        if ((rc = DISPATCH(c, run_metropolis_learned_sweep, c, p, c->counts + i))) return rc;
        if (plaq && (rc = DISPATCH(c, run_plaquette, c, c->field[GLACX_FIELD_MAIN], c->results + i))) return rc;
    }
    std::vector<unsigned long long> counts(n_sweeps);
    HIP_TRY(c, hipMemcpyAsync(counts.data(), c->counts, sizeof(unsigned long long) * n_sweeps, hipMemcpyDeviceToHost,
                              c->stream));
    if (plaq) HIP_TRY(c, hipMemcpyAsync(plaq, c->results, sizeof(double) * n_sweeps, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (accept)
        for (int i = 0; i < n_sweeps; i++) accept[i] = (double)counts[i] / (4.0 * n_hit * (double)c->geom.v);
    return GLACX_OK;
}

int glacx_timing_enable(glacx_ctx *c, int on)
{
    CTX_CHECK(c);
    c->timing = on != 0;
    return GLACX_OK;
}

int glacx_timing_reset(glacx_ctx *c)
{
    CTX_CHECK(c);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->timed.clear();
    c->ev_used = 0;
    for (int k = 0; k < GLACX_K_COUNT; k++) {
        c->t_launches[k] = 0;
        c->t_ms[k] = 0.0;
    }
    return GLACX_OK;
}

int glacx_timing_read(glacx_ctx *c, int kernel, long long *launches, double *total_ms)
{
    CTX_CHECK(c);
    if (kernel < 0 || kernel >= GLACX_K_COUNT) return fail(c, GLACX_E_ARG, "timing_read: bad kernel id");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (auto &t : c->timed) {
        float ms = 0.f;
        HIP_TRY(c, hipEventElapsedTime(&ms, t.second.first, t.second.second));
        c->t_launches[t.first]++;
        c->t_ms[t.first] += ms;
    }
    c->timed.clear();
    c->ev_used = 0;
    if (launches) *launches = c->t_launches[kernel];
    if (total_ms) *total_ms = c->t_ms[kernel];
    return GLACX_OK;
}

}  // extern "C"
