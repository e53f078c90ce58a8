// tp_flow_engine.cpp -- C ABI of the learned proposal (include/tp.h, "Learned
// proposal"): flow parameters, reverse-KL training on the GPU, sampling and
// density queries, and the weighted event loop.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "tp.h"
#include "tp_ctx.h"
#include "tp_flow.h"
#include "tp_kernels.h"

static_assert(tpf::NPARAM == TP_FLOW_NPARAM, "parameter count of the ABI and the kernels");

namespace {

using namespace tpe;

struct Flow {
    Ctx *c = nullptr;
    DevBuf w, wbf, m, v, grad;            // parameters, bf16 copy, Adam moments, gradient
    DevBuf zsave, save, u, logq, gz, logf; // training batch
    DevBuf gpart;                         // per-block partial weight gradients
    DevBuf f, loss, vin, io;              // integrand values, per-step loss, staging
    DevBuf ru, rinvq;                     // weighted-run chunk
    int step = 0;                         // Adam steps taken
};

std::string *flow_err(void *flow)
{
    Flow *f = static_cast<Flow *>(flow);
    return f && f->c ? &f->c->err : nullptr;
}

// Batches per weighted-run chunk: two full strides of the 3072-wavefront
// event-loop grid (6.1e7 events: 5.9 GB of u).
constexpr uint64_t RUN_CHUNK_BATCHES = 6144;
constexpr uint64_t UNIFORM_KEY_XOR = 0x5555555555555555ull;

int ffail(Flow *f, int code, const std::string &msg) { return fail(f ? f->c : nullptr, code, msg); }

#define FHIP(fl, expr) HIPCHK((fl)->c, expr)

uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Host initialisation: hidden layers uniform in +-1/sqrt(fan_in) (the
// usual default for linear layers), output layer zero so that the initial
// flow is the identity.
void init_params(uint64_t seed, std::vector<float> &w)
{
    w.assign(tpf::NPARAM, 0.f);
    uint64_t k = 0;
    auto uni = [&](float bound) {
        const double x = (double)(mix64(seed + (++k) * 0x9E3779B97F4A7C15ull) >> 11) * 0x1.0p-53;
        return (float)((2. * x - 1.) * bound);
    };
    for (int l = 0; l < tpf::L; ++l) {
        float *p = w.data() + (size_t)l * tpf::LP;
        const float b1 = 1.f / std::sqrt((float)tpf::D), b2 = 1.f / std::sqrt((float)tpf::H);
        for (int r = 0; r < tpf::H; ++r) {
            for (int cc = 0; cc < tpf::D; ++cc)
                p[tpf::OFF_W1 + r * tpf::DP + cc] = uni(b1);
            p[tpf::OFF_B1 + r] = uni(b1);
        }
        for (int r = 0; r < tpf::H; ++r) {
            for (int cc = 0; cc < tpf::H; ++cc)
                p[tpf::OFF_W2 + r * tpf::H + cc] = uni(b2);
            p[tpf::OFF_B2 + r] = uni(b2);
        }
    }
}

// Zero the padding entries (inputs 12..15 of W1, rows 12..15 and 28..31 of
// W3 and b3) so that they cannot leak into the outputs.
void zero_padding(std::vector<float> &w)
{
    for (int l = 0; l < tpf::L; ++l) {
        float *p = w.data() + (size_t)l * tpf::LP;
        for (int r = 0; r < tpf::H; ++r)
            for (int cc = tpf::D; cc < tpf::DP; ++cc)
                p[tpf::OFF_W1 + r * tpf::DP + cc] = 0.f;
        for (int r = 0; r < tpf::O; ++r)
            if ((r >= tpf::D && r < 16) || r >= 16 + tpf::D) {
                for (int cc = 0; cc < tpf::H; ++cc)
                    p[tpf::OFF_W3 + r * tpf::H + cc] = 0.f;
                p[tpf::OFF_B3 + r] = 0.f;
            }
    }
}

int upload_params(Flow *f, std::vector<float> w)
{
    zero_padding(w);
    const size_t bytes = (size_t)tpf::NPARAM * sizeof(float);
    int rc = ensure(f->c, f->w, bytes);
    if (!rc)
        rc = ensure(f->c, f->wbf, (size_t)tpf::NPARAM * sizeof(uint16_t));
    if (!rc)
        rc = ensure(f->c, f->m, bytes);
    if (!rc)
        rc = ensure(f->c, f->v, bytes);
    if (!rc)
        rc = ensure(f->c, f->grad, bytes);
    if (rc)
        return rc;
    FHIP(f, hipMemcpyAsync(f->w.p, w.data(), bytes, hipMemcpyHostToDevice, f->c->stream));
    FHIP(f, hipMemsetAsync(f->m.p, 0, bytes, f->c->stream));
    FHIP(f, hipMemsetAsync(f->v.p, 0, bytes, f->c->stream));
    tpf::launch_flow_to_bf16(static_cast<const float *>(f->w.p), static_cast<uint16_t *>(f->wbf.p), f->c->stream);
    FHIP(f, hipGetLastError());
    FHIP(f, hipStreamSynchronize(f->c->stream)); // w is a host temporary
    f->step = 0;
    return TP_OK;
}

tpf::FlowKernelArgs base_args(const Flow *f, uint64_t n)
{
    tpf::FlowKernelArgs a;
    std::memset(&a, 0, sizeof a);
    a.wbf = static_cast<const uint16_t *>(f->wbf.p);
    a.n = n;
    a.alpha = 1.;
    a.max_blocks = (uint32_t)f->c->cus;
    return a;
}

// Device pointer for a caller buffer: itself when on_device, else a staging
// buffer (region `slot` of f->io, `bytes` long).
struct Stage {
    void *dev;
    void *host;
    size_t bytes;
};

int stage_layout(Flow *f, std::vector<Stage> &st, int on_device)
{
    if (on_device)
        return TP_OK;
    size_t total = 0;
    for (const Stage &s : st)
        if (s.host)
            total += (s.bytes + 255) / 256 * 256;
    int rc = ensure(f->c, f->io, std::max<size_t>(total, 256));
    if (rc)
        return rc;
    size_t off = 0;
    for (Stage &s : st)
        if (s.host) {
            s.dev = static_cast<char *>(f->io.p) + off;
            off += (s.bytes + 255) / 256 * 256;
        }
    return TP_OK;
}

int check_params(Flow *f, const tp_params *p)
{
    if (!p)
        return ffail(f, TP_ERR_ARG, "null params");
    if (!(p->e_total > 0.) || !std::isfinite(p->e_total))
        return ffail(f, TP_ERR_ARG, "e_total must be positive and finite");
    return TP_OK;
}

} // namespace

extern "C" int tp_flow_default_opts(tp_flow_opts *o)
{
    return tp_guard(nullptr, [&]() -> int {
        if (!o)
            return TP_ERR_ARG;
        o->alpha = 0.7;
        o->batch = 65536;
        o->steps = 2000;
        o->lr = 3e-3f;
        o->beta1 = 0.9f;
        o->beta2 = 0.999f;
        o->adam_eps = 1e-8f;
        o->grad_clip = 100.f;
        o->eps_rel = 1e-3;
        o->cut_width = 0.05;
        o->target = 2;
        o->cap_rel = 10.;
        o->seed = 1;
        return TP_OK;
    });
}

extern "C" int tp_flow_create(void *ctx, uint64_t init_seed, void **flow_out)
{
    return tp_guard(ctx_err(ctx), [&]() -> int {
        Ctx *c = static_cast<Ctx *>(ctx);
        if (!c || !flow_out)
            return TP_ERR_ARG;
        *flow_out = nullptr;
        HIPCHK(c, hipSetDevice(c->device));
        Flow *f = new Flow;
        f->c = c;
        std::vector<float> w;
        init_params(init_seed, w);
        int rc = upload_params(f, w);
        if (rc) {
            tp_flow_destroy(f);
            return rc;
        }
        *flow_out = f;
        return TP_OK;
    });
}

extern "C" int tp_flow_destroy(void *flow)
{
    return tp_guard(nullptr, [&]() -> int {
        Flow *f = static_cast<Flow *>(flow);
        if (!f)
            return TP_ERR_ARG;
        (void)hipSetDevice(f->c->device);
        (void)hipStreamSynchronize(f->c->stream);
        for (DevBuf *b : {&f->w, &f->wbf, &f->m, &f->v, &f->grad, &f->zsave, &f->save, &f->u, &f->logq, &f->gz, &f->logf,
                          &f->gpart, &f->f, &f->loss, &f->vin, &f->io, &f->ru, &f->rinvq})
            if (b->p)
                (void)hipFree(b->p);
        delete f;
        return TP_OK;
    });
}

extern "C" int tp_flow_get_params(void *flow, float *w)
{
    return tp_guard(flow_err(flow), [&]() -> int {
        Flow *f = static_cast<Flow *>(flow);
        if (!f || !w)
            return TP_ERR_ARG;
        FHIP(f, hipSetDevice(f->c->device));
        FHIP(f, hipMemcpyAsync(w, f->w.p, (size_t)tpf::NPARAM * sizeof(float), hipMemcpyDeviceToHost, f->c->stream));
        FHIP(f, hipStreamSynchronize(f->c->stream));
        return TP_OK;
    });
}

extern "C" int tp_flow_set_params(void *flow, const float *w)
{
    return tp_guard(flow_err(flow), [&]() -> int {
        Flow *f = static_cast<Flow *>(flow);
        if (!f || !w)
            return TP_ERR_ARG;
        FHIP(f, hipSetDevice(f->c->device));
        return upload_params(f, std::vector<float>(w, w + tpf::NPARAM));
    });
}

extern "C" int tp_flow_sample(void *flow, uint64_t key, uint64_t sample0, uint64_t n, const float *v_in,
                              int on_device, double *u, float *z, float *logq)
{
    return tp_guard(flow_err(flow), [&]() -> int {
        Flow *f = static_cast<Flow *>(flow);
        if (!f)
            return TP_ERR_ARG;
        if (n == 0)
            return TP_OK;
        FHIP(f, hipSetDevice(f->c->device));
        std::vector<Stage> st = {{(void *)v_in, (void *)v_in, v_in ? n * 16 * sizeof(float) : 0},
                                 {u, u, u ? n * 12 * sizeof(double) : 0},
                                 {z, z, z ? n * 12 * sizeof(float) : 0},
                                 {logq, logq, logq ? n * sizeof(float) : 0}};
        int rc = stage_layout(f, st, on_device);
        if (rc)
            return rc;
        if (!on_device && v_in)
            FHIP(f, hipMemcpyAsync(st[0].dev, v_in, st[0].bytes, hipMemcpyHostToDevice, f->c->stream));
        tpf::FlowKernelArgs a = base_args(f, n);
        a.v_in = static_cast<const float *>(st[0].dev);
        a.key = key;
        a.sample0 = sample0;
        a.u_out = static_cast<double *>(st[1].dev);
        a.z_out = static_cast<float *>(st[2].dev);
        a.logq_out = static_cast<float *>(st[3].dev);
        tpf::launch_flow_forward(a, f->c->stream);
        FHIP(f, hipGetLastError());
        if (!on_device) {
            for (int i = 1; i < 4; ++i)
                if (st[i].host)
                    FHIP(f, hipMemcpyAsync(st[i].host, st[i].dev, st[i].bytes, hipMemcpyDeviceToHost, f->c->stream));
            FHIP(f, hipStreamSynchronize(f->c->stream));
        }
        return TP_OK;
    });
}

extern "C" int tp_flow_density(void *flow, uint64_t n, const double *u, int on_device, float *logq)
{
    return tp_guard(flow_err(flow), [&]() -> int {
        Flow *f = static_cast<Flow *>(flow);
        if (!f)
            return TP_ERR_ARG;
        if (!u || !logq)
            return ffail(f, TP_ERR_ARG, "tp_flow_density: null u or logq");
        if (n == 0)
            return TP_OK;
        FHIP(f, hipSetDevice(f->c->device));
        std::vector<Stage> st = {{(void *)u, (void *)u, n * 12 * sizeof(double)}, {logq, logq, n * sizeof(float)}};
        int rc = stage_layout(f, st, on_device);
        if (rc)
            return rc;
        if (!on_device)
            FHIP(f, hipMemcpyAsync(st[0].dev, u, st[0].bytes, hipMemcpyHostToDevice, f->c->stream));
        tpf::FlowKernelArgs a = base_args(f, n);
        a.u_in = static_cast<const double *>(st[0].dev);
        a.logq_out = static_cast<float *>(st[1].dev);
        tpf::launch_flow_inverse(a, f->c->stream);
        FHIP(f, hipGetLastError());
        if (!on_device) {
            FHIP(f, hipMemcpyAsync(logq, st[1].dev, st[1].bytes, hipMemcpyDeviceToHost, f->c->stream));
            FHIP(f, hipStreamSynchronize(f->c->stream));
        }
        return TP_OK;
    });
}

namespace {

int ensure_train(Flow *f, uint64_t n)
{
    int rc = ensure(f->c, f->zsave, (size_t)tpf::L * n * 16 * sizeof(float));
    if (!rc)
        rc = ensure(f->c, f->save, (size_t)tpf::L * tpf::SAVE_F * n * sizeof(uint16_t));
    if (!rc)
        rc = ensure(f->c, f->u, n * 12 * sizeof(double));
    if (!rc)
        rc = ensure(f->c, f->logq, n * sizeof(float));
    if (!rc)
        rc = ensure(f->c, f->gz, n * 16 * sizeof(float));
    if (!rc)
        rc = ensure(f->c, f->logf, n * sizeof(float));
    if (!rc)
        rc = ensure(f->c, f->f, n * sizeof(double));
    if (!rc)
        rc = ensure(f->c, f->gpart, tpf::flow_wgrad_part_bytes(n));
    return rc;
}

// Backward pass and batch weight gradients for zsave (filled by a forward
// pass with zsave set) and gz in f->gz.
int enqueue_grad(Flow *f, uint64_t n, float *gv)
{
    tpf::FlowBackArgs b;
    std::memset(&b, 0, sizeof b);
    b.wbf = static_cast<const uint16_t *>(f->wbf.p);
    b.n = n;
    b.zsave = static_cast<const float *>(f->zsave.p);
    b.gz = static_cast<const float *>(f->gz.p);
    b.save = static_cast<uint16_t *>(f->save.p);
    b.gv = gv;
    tpf::launch_flow_backward(b, f->c->stream);
    FHIP(f, hipGetLastError());
    tpf::launch_flow_wgrad(static_cast<const uint16_t *>(f->save.p), n, 1.f / (float)n, static_cast<float *>(f->gpart.p),
                           static_cast<float *>(f->grad.p), f->c->stream);
    FHIP(f, hipGetLastError());
    return TP_OK;
}

} // namespace

extern "C" int tp_flow_grad(void *flow, uint64_t n, const float *v_in, const float *gz, int on_device, float *grad,
                            float *gv)
{
    return tp_guard(flow_err(flow), [&]() -> int {
        Flow *f = static_cast<Flow *>(flow);
        if (!f)
            return TP_ERR_ARG;
        if (!v_in || !gz || !grad)
            return ffail(f, TP_ERR_ARG, "tp_flow_grad: null v_in, gz or grad");
        if (n == 0 || n % 16)
            return ffail(f, TP_ERR_ARG, "tp_flow_grad: n must be a positive multiple of 16");
        FHIP(f, hipSetDevice(f->c->device));
        int rc = ensure_train(f, n);
        if (rc)
            return rc;
        std::vector<Stage> st = {{(void *)v_in, (void *)v_in, n * 16 * sizeof(float)},
                                 {gv, gv, gv ? n * 16 * sizeof(float) : 0}};
        rc = stage_layout(f, st, on_device);
        if (rc)
            return rc;
        if (!on_device)
            FHIP(f, hipMemcpyAsync(st[0].dev, v_in, st[0].bytes, hipMemcpyHostToDevice, f->c->stream));
        FHIP(f, hipMemcpyAsync(f->gz.p, gz, n * 16 * sizeof(float),
                               on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, f->c->stream));
        tpf::FlowKernelArgs a = base_args(f, n);
        a.v_in = static_cast<const float *>(st[0].dev);
        a.zsave = static_cast<float *>(f->zsave.p);
        tpf::launch_flow_forward(a, f->c->stream);
        FHIP(f, hipGetLastError());
        rc = enqueue_grad(f, n, static_cast<float *>(st[1].dev));
        if (rc)
            return rc;
        FHIP(f, hipMemcpyAsync(grad, f->grad.p, (size_t)tpf::NPARAM * sizeof(float),
                               on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, f->c->stream));
        if (!on_device && gv)
            FHIP(f, hipMemcpyAsync(gv, st[1].dev, st[1].bytes, hipMemcpyDeviceToHost, f->c->stream));
        FHIP(f, hipStreamSynchronize(f->c->stream));
        return TP_OK;
    });
}

extern "C" int tp_flow_integrand(void *flow, const tp_params *params, uint64_t n, const double *u, double eps,
                                 double tau, double fcap, int target, float clip, int on_device, double *fv,
                                 float *gz, float *logf)
{
    return tp_guard(flow_err(flow), [&]() -> int {
        Flow *f = static_cast<Flow *>(flow);
        if (!f)
            return TP_ERR_ARG;
        int rc = check_params(f, params);
        if (rc)
            return rc;
        if (!u || !gz || !logf)
            return ffail(f, TP_ERR_ARG, "tp_flow_integrand: null u, gz or logf");
        if (n == 0)
            return TP_OK;
        FHIP(f, hipSetDevice(f->c->device));
        std::vector<Stage> st = {{(void *)u, (void *)u, n * 12 * sizeof(double)},
                                 {fv, fv, fv ? n * sizeof(double) : 0},
                                 {gz, gz, n * 16 * sizeof(float)},
                                 {logf, logf, n * sizeof(float)}};
        rc = stage_layout(f, st, on_device);
        if (rc)
            return rc;
        if (!on_device)
            FHIP(f, hipMemcpyAsync(st[0].dev, u, st[0].bytes, hipMemcpyHostToDevice, f->c->stream));
        const tpd::Params P = make_params(params);
        if (target < 0 || target > 2 || (target > 0 && !(tau > 0. && tau <= 0.1)))
            return ffail(f, TP_ERR_ARG, "tp_flow_integrand: target must be 0, 1 or 2, with tau in (0, 0.1] for 1 and 2");
        tpf::launch_flow_integrand(P, params->sigma_contribs, n, static_cast<const double *>(st[0].dev), eps, tau, fcap,
                                   target, clip,
                                   static_cast<float *>(st[2].dev), static_cast<float *>(st[3].dev),
                                   static_cast<double *>(st[1].dev), f->c->stream);
        FHIP(f, hipGetLastError());
        if (!on_device) {
            for (int i = 1; i < 4; ++i)
                if (st[i].host)
                    FHIP(f, hipMemcpyAsync(st[i].host, st[i].dev, st[i].bytes, hipMemcpyDeviceToHost, f->c->stream));
            FHIP(f, hipStreamSynchronize(f->c->stream));
        }
        return TP_OK;
    });
}

extern "C" int tp_flow_train(void *flow, const tp_params *params, const tp_flow_opts *o, float *loss, tp_stats *stats)
{
    return tp_guard(flow_err(flow), [&]() -> int {
        Flow *f = static_cast<Flow *>(flow);
        if (!f)
            return TP_ERR_ARG;
        int rc = check_params(f, params);
        if (rc)
            return rc;
        if (!o)
            return ffail(f, TP_ERR_ARG, "tp_flow_train: null options");
        if (o->batch == 0 || o->batch % 16 || o->steps < 0)
            return ffail(f, TP_ERR_ARG, "tp_flow_train: batch must be a positive multiple of 16, steps >= 0");
        if (!(o->eps_rel > 0.) || !(o->grad_clip > 0.f) || !(o->lr > 0.f))
            return ffail(f, TP_ERR_ARG, "tp_flow_train: eps_rel, grad_clip and lr must be positive");
        if (o->target < 0 || o->target > 2 || (o->target > 0 && !(o->cut_width > 0. && o->cut_width <= 0.1)))
            return ffail(f, TP_ERR_ARG, "tp_flow_train: target must be 0, 1 or 2, with cut_width in (0, 0.1] for 1 and 2");
        if (o->target == 2 && !(o->cap_rel > 0.))
            return ffail(f, TP_ERR_ARG, "tp_flow_train: cap_rel must be positive");
        Ctx *c = f->c;
        FHIP(f, hipSetDevice(c->device));
        const uint64_t n = o->batch;
        rc = ensure_train(f, n);
        if (!rc)
            rc = ensure(c, f->loss, std::max<size_t>(1, (size_t)o->steps) * sizeof(float));
        if (rc)
            return rc;
        const tpd::Params P = make_params(params);
        HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
        // eps from the mean of f over the cube (uniform points of the same stream
        // family, a key of their own).
        tpf::FlowKernelArgs a = base_args(f, n);
        a.key = o->seed ^ UNIFORM_KEY_XOR;
        a.u_out = static_cast<double *>(f->u.p);
        tpf::launch_flow_inverse(a, c->stream);
        FHIP(f, hipGetLastError());
        tpf::launch_flow_integrand(P, params->sigma_contribs, n, static_cast<const double *>(f->u.p), 0., 0., 0., 0,
                                   o->grad_clip,
                                   static_cast<float *>(f->gz.p), static_cast<float *>(f->logf.p),
                                   static_cast<double *>(f->f.p), c->stream);
        FHIP(f, hipGetLastError());
        std::vector<double> fh(n);
        FHIP(f, hipMemcpyAsync(fh.data(), f->f.p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        FHIP(f, hipStreamSynchronize(c->stream));
        double mean_f = 0.;
        for (double x : fh)
            mean_f += x;
        mean_f /= (double)n;
        if (!(mean_f > 0.) || !std::isfinite(mean_f))
            return ffail(f, TP_ERR_ARG, "tp_flow_train: the integrand has no positive mean over the cube");
        const double eps = o->eps_rel * mean_f;
        HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
        for (int s = 0; s < o->steps; ++s) {
            tpf::FlowKernelArgs fa = base_args(f, n);
            fa.key = o->seed;
            fa.sample0 = (uint64_t)s * n;
            fa.u_out = static_cast<double *>(f->u.p);
            fa.logq_out = static_cast<float *>(f->logq.p);
            fa.zsave = static_cast<float *>(f->zsave.p);
            tpf::launch_flow_forward(fa, c->stream);
            FHIP(f, hipGetLastError());
            tpf::launch_flow_integrand(P, params->sigma_contribs, n, static_cast<const double *>(f->u.p), eps,
                                       o->cut_width, o->cap_rel * mean_f, o->target, o->grad_clip, static_cast<float *>(f->gz.p), static_cast<float *>(f->logf.p),
                                       nullptr, c->stream);
            FHIP(f, hipGetLastError());
            tpf::launch_flow_loss(static_cast<const float *>(f->logq.p), static_cast<const float *>(f->logf.p), n,
                                  static_cast<float *>(f->loss.p) + s, c->stream);
            FHIP(f, hipGetLastError());
            rc = enqueue_grad(f, n, nullptr);
            if (rc)
                return rc;
            ++f->step;
            tpf::launch_flow_adam(static_cast<float *>(f->w.p), static_cast<uint16_t *>(f->wbf.p),
                                  static_cast<float *>(f->m.p), static_cast<float *>(f->v.p),
                                  static_cast<const float *>(f->grad.p), f->step, o->lr, o->beta1, o->beta2, o->adam_eps,
                                  c->stream);
            FHIP(f, hipGetLastError());
        }
        HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
        HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
        if (loss && o->steps > 0)
            FHIP(f, hipMemcpyAsync(loss, f->loss.p, (size_t)o->steps * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        FHIP(f, hipStreamSynchronize(c->stream));
        tp_stats t{};
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev[0], c->ev[3]) == hipSuccess)
            t.device_ms = ms;
        if (hipEventElapsedTime(&ms, c->ev[1], c->ev[2]) == hipSuccess)
            t.kernel_ms = ms;
        t.events = (uint64_t)o->steps * n;
        t.batches = (uint64_t)o->steps;
        c->last = t;
        c->timed = false;
        if (stats)
            *stats = t;
        return TP_OK;
    });
}

extern "C" int tp_flow_run_device(void *flow, const tp_params *params, double alpha, uint64_t key,
                                  uint64_t num_events, tp_acc *per_batch_dev, tp_acc *merged_dev, tp_stats *stats)
{
    return tp_guard(flow_err(flow), [&]() -> int {
        Flow *f = static_cast<Flow *>(flow);
        if (!f)
            return TP_ERR_ARG;
        int rc = check_params(f, params);
        if (rc)
            return rc;
        if (!(alpha >= 0. && alpha <= 1.))
            return ffail(f, TP_ERR_ARG, "tp_flow_run_device: alpha must lie in [0, 1]");
        if (num_events == 0)
            return ffail(f, TP_ERR_ARG, "Must simulate at least one event");
        if (!per_batch_dev)
            return ffail(f, TP_ERR_ARG, "null per-batch output");
        Ctx *c = f->c;
        FHIP(f, hipSetDevice(c->device));
        const uint32_t bs = TP_BATCH_SIZE;
        const uint64_t nb = (num_events + bs - 1) / bs;
        const uint64_t chunk_b = std::min<uint64_t>(nb, RUN_CHUNK_BATCHES);
        rc = ensure(c, f->ru, chunk_b * bs * 12 * sizeof(double));
        if (!rc)
            rc = ensure(c, f->rinvq, chunk_b * bs * sizeof(double));
        if (rc)
            return rc;
        tpk::FusedArgs fa;
        std::memset(&fa, 0, sizeof fa);
        fa.P = make_params(params);
        fa.num_events = num_events;
        fa.batch_size = bs;
        fa.precision = TP_PREC_F64;
        // u as 12 rows of the chunk's events (flow samples first, then the
        // uniform points), the layout the event loop loads contiguously
        const uint64_t ustride = chunk_b * bs;
        fa.u = static_cast<const double *>(f->ru.p);
        fa.u_stride = ustride;
        fa.invq = static_cast<const double *>(f->rinvq.p);
        HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
        HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
        uint64_t flow_done = 0, unif_done = 0;
        for (uint64_t b0 = 0; b0 < nb; b0 += chunk_b) {
            const uint64_t b1 = std::min(nb, b0 + chunk_b);
            const uint64_t e0 = b0 * bs, e1 = std::min<uint64_t>(b1 * bs, num_events);
            const uint64_t ne = e1 - e0;
            const uint64_t nf = (uint64_t)std::llround(alpha * (double)ne);
            if (nf > 0) {
                tpf::FlowKernelArgs a = base_args(f, nf);
                a.key = key;
                a.sample0 = flow_done;
                a.u_out = static_cast<double *>(f->ru.p);
                a.u_stride = ustride;
                a.alpha = alpha;
                a.invq_out = static_cast<double *>(f->rinvq.p);
                tpf::launch_flow_forward(a, c->stream);
                FHIP(f, hipGetLastError());
            }
            if (ne > nf) {
                tpf::FlowKernelArgs a = base_args(f, ne - nf);
                a.key = key ^ UNIFORM_KEY_XOR;
                a.sample0 = unif_done;
                a.u_out = static_cast<double *>(f->ru.p) + nf;
                a.u_stride = ustride;
                a.alpha = alpha;
                a.invq_out = static_cast<double *>(f->rinvq.p) + nf;
                tpf::launch_flow_inverse(a, c->stream);
                FHIP(f, hipGetLastError());
            }
            flow_done += nf;
            unif_done += ne - nf;
            fa.nb = b1 - b0;
            fa.b_first = b0;
            fa.out = per_batch_dev + b0;
            tpk::launch_fused(fa, 2, grid_for(c, 2, 0, fa.nb), c->stream);
            FHIP(f, hipGetLastError());
        }
        HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
        if (merged_dev) {
            tpk::launch_merge(per_batch_dev, nb, merged_dev, c->stream);
            FHIP(f, hipGetLastError());
        }
        HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
        c->last = tp_stats{};
        c->last.events = num_events;
        c->last.batches = nb;
        c->timed = true;
        if (stats)
            *stats = c->last;
        return TP_OK;
    });
}

extern "C" int tp_flow_run(void *flow, const tp_params *params, double alpha, uint64_t key, uint64_t num_events,
                           tp_acc *per_batch_out, tp_acc *merged_out, tp_stats *stats)
{
    return tp_guard(flow_err(flow), [&]() -> int {
        Flow *f = static_cast<Flow *>(flow);
        if (!f)
            return TP_ERR_ARG;
        if (!per_batch_out)
            return ffail(f, TP_ERR_ARG, "null per-batch output");
        if (num_events == 0)
            return ffail(f, TP_ERR_ARG, "Must simulate at least one event");
        Ctx *c = f->c;
        FHIP(f, hipSetDevice(c->device));
        const uint64_t nb = (num_events + TP_BATCH_SIZE - 1) / TP_BATCH_SIZE;
        int rc = ensure(c, c->per_batch, nb * sizeof(tp_acc));
        if (!rc)
            rc = ensure(c, c->merged, sizeof(tp_acc));
        if (!rc)
            rc = tp_flow_run_device(flow, params, alpha, key, num_events, static_cast<tp_acc *>(c->per_batch.p),
                                    merged_out ? static_cast<tp_acc *>(c->merged.p) : nullptr, nullptr);
        if (rc)
            return rc;
        FHIP(f, hipMemcpyAsync(per_batch_out, c->per_batch.p, nb * sizeof(tp_acc), hipMemcpyDeviceToHost, c->stream));
        if (merged_out)
            FHIP(f, hipMemcpyAsync(merged_out, c->merged.p, sizeof(tp_acc), hipMemcpyDeviceToHost, c->stream));
        return tp_sync(c, stats);
    });
}
