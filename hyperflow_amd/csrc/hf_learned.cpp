// hf_learned.cpp -- C ABI of the learned coarse-grid correction
// (include/hyperflow_learned.h): network state, training step, and the
// coupling to the solver context.
//
// Layers (real channels -> storage channels of the NHWC images):
//   1: 4 -> 32   (input stored with 8 channels)
//   2-4: 32 -> 32
//   5: 32 -> 4   (output conv computed with 8 output rows, 4 real)
// Parameters: FP32 master copy W[9][ci][co] + b[co] per layer, Adam moments,
// and two BF16 packings of every W: the forward A operands and (layers 2-5)
// the data-gradient A operands.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/hyperflow_learned.h"
#include "hf_cnn.h"
#include "hf_context.h"

namespace {

struct Layer {
    int ci, co;     // real channels
    int cis, cos;   // storage channels
    int64_t off;    // parameter offset (W, then b)
};
constexpr int NL = 5;
constexpr int64_t P1 = 9 * 4 * 32 + 32, PH = 9 * 32 * 32 + 32, P5 = 9 * 32 * 4 + 4;
const Layer LAYERS[NL] = {
    {4, 32, 8, 32, 0},
    {32, 32, 32, 32, P1},
    {32, 32, 32, 32, P1 + PH},
    {32, 32, 32, 32, P1 + 2 * PH},
    {32, 4, 32, 8, P1 + 3 * PH},
};
constexpr int64_t NPARAM = P1 + 3 * PH + P5;
constexpr int CONV_ROWS = 8;    // rows per wave of the convolutions
constexpr int WGRAD_ROWS = 16;  // rows per workgroup of the weight gradient

int fail(int code, const std::string& msg) { return hf_fail(code, msg); }

}  // namespace

struct hf_net {
    hf_net_desc d;
    hipStream_t stream = nullptr;
    float* params = nullptr;
    float* grads = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    int64_t step = 0;
    uint16_t* wf[NL] = {};
    uint16_t* wb[NL] = {};
    // image buffers for one (nx, ny)
    int nx = 0, ny = 0;
    uint16_t* x = nullptr;
    uint16_t* act[4] = {};
    uint16_t* g8 = nullptr;
    uint16_t* gz[2] = {};
    float* y = nullptr;
    uint8_t* shock = nullptr;
    float* part = nullptr;
    float* part2 = nullptr;   // second stage of the gradient reduction
    float* lossp = nullptr;
    float* loss = nullptr;
    uint32_t* fallback = nullptr;
    // ordering between the net's stream and solver streams
    hipEvent_t done = nullptr;
    hipStream_t done_stream = nullptr;
};

namespace {

void free_images(hf_net* n)
{
    (void)hipFree(n->x);
    for (auto& p : n->act) (void)hipFree(p), p = nullptr;
    (void)hipFree(n->g8);
    for (auto& p : n->gz) (void)hipFree(p), p = nullptr;
    (void)hipFree(n->y);
    (void)hipFree(n->shock);
    (void)hipFree(n->part);
    (void)hipFree(n->part2);
    (void)hipFree(n->lossp);
    n->x = n->g8 = nullptr;
    n->y = n->part = n->part2 = n->lossp = nullptr;
    n->shock = nullptr;
    n->nx = n->ny = 0;
}

// Buffers for nx x ny images; borders zeroed once.
int ensure_images(hf_net* n, int nx, int ny)
{
    if (nx < 64 || nx % 64 != 0 || ny < 1) return fail(HF_EINVAL, "learned correction: nx must be a positive multiple of 64");
    if (n->nx == nx && n->ny == ny) return HF_OK;
    HF_HIP(hipDeviceSynchronize());
    free_images(n);
    const size_t pix = (size_t)(nx + 2) * (ny + 2);
    HF_HIP(hipMalloc(&n->x, pix * 8 * 2));
    HF_HIP(hipMalloc(&n->g8, pix * 8 * 2));
    HF_HIP(hipMemset(n->x, 0, pix * 8 * 2));
    HF_HIP(hipMemset(n->g8, 0, pix * 8 * 2));
    for (auto& p : n->act) {
        HF_HIP(hipMalloc(&p, pix * 32 * 2));
        HF_HIP(hipMemset(p, 0, pix * 32 * 2));
    }
    for (auto& p : n->gz) {
        HF_HIP(hipMalloc(&p, pix * 32 * 2));
        HF_HIP(hipMemset(p, 0, pix * 32 * 2));
    }
    HF_HIP(hipMalloc(&n->y, (size_t)nx * ny * 4 * sizeof(float)));
    HF_HIP(hipMalloc(&n->shock, (size_t)nx * ny));
    const int nwg = hfn::wgrad_workgroups(nx, ny, WGRAD_ROWS);
    HF_HIP(hipMalloc(&n->part, (size_t)nwg * (9 * 32 * 32 + 32) * sizeof(float)));
    HF_HIP(hipMalloc(&n->part2, (size_t)hfn::REDUCE_SPLIT * (9 * 32 * 32 + 32) * sizeof(float)));
    const int nloss = (int)(((int64_t)nx * ny + 255) / 256);
    HF_HIP(hipMalloc(&n->lossp, (size_t)nloss * sizeof(float)));
    n->nx = nx;
    n->ny = ny;
    return HF_OK;
}

// Makes stream s wait for the last network work if it ran on another stream.
int order_after(hf_net* n, hipStream_t s)
{
    if (n->done_stream != s && n->done_stream != (hipStream_t)-1) HF_HIP(hipStreamWaitEvent(s, n->done, 0));
    return HF_OK;
}
int mark_done(hf_net* n, hipStream_t s)
{
    HF_HIP(hipEventRecord(n->done, s));
    n->done_stream = s;
    return HF_OK;
}

int pack_weights(hf_net* n, hipStream_t s)
{
    for (int l = 0; l < NL; l++) {
        const Layer& L = LAYERS[l];
        HF_HIP(hfn::launch_pack_weights(L.cis, L.cos, n->params + L.off, L.ci, L.co, 0, n->wf[l], s));
        if (l > 0) HF_HIP(hfn::launch_pack_weights(L.cos, L.cis, n->params + L.off, L.ci, L.co, 1, n->wb[l], s));
    }
    return HF_OK;
}

hfn::ConvArgs conv_args(const hf_net* n, int l, const uint16_t* in, int nx, int ny)
{
    hfn::ConvArgs a;
    std::memset(&a, 0, sizeof a);
    a.in = in;
    a.wa = n->wf[l];
    a.bias = n->params + LAYERS[l].off + 9 * LAYERS[l].ci * LAYERS[l].co;
    a.co_real = LAYERS[l].co;
    a.nx = nx;
    a.ny = ny;
    a.rows = CONV_ROWS;
    return a;
}

// Input packing and layers 1-4 (activations act[0..3]).
int forward_hidden(hf_net* n, const float* U, int64_t plane, int pitch, int nx, int ny, bool sensor, hipStream_t s)
{
    hfn::PackArgs p;
    std::memset(&p, 0, sizeof p);
    p.U = U;
    p.plane = plane;
    p.pitch = pitch;
    p.nx = nx;
    p.ny = ny;
    for (int r = 0; r < 4; r++) {
        p.mean[r] = n->d.in_mean[r];
        p.iscale[r] = n->d.in_scale[r];
    }
    p.X = n->x;
    p.shock = sensor ? n->shock : nullptr;
    p.gamma = n->d.gamma;
    p.theta = n->d.shock_threshold;
    HF_HIP(hfn::launch_pack(p, s));
    const uint16_t* in = n->x;
    for (int l = 0; l < 4; l++) {
        hfn::ConvArgs a = conv_args(n, l, in, nx, ny);
        a.out = n->act[l];
        HF_HIP(hfn::launch_conv(LAYERS[l].cis, LAYERS[l].cos, hfn::CONV_HIDDEN, a, s));
        in = n->act[l];
    }
    return HF_OK;
}

int wgrad(hf_net* n, int l, const uint16_t* a_in, const uint16_t* gz, int nx, int ny, hipStream_t s)
{
    const Layer& L = LAYERS[l];
    hfn::WgradArgs w;
    w.a = a_in;
    w.gz = gz;
    w.part = n->part;
    w.nx = nx;
    w.ny = ny;
    w.rows = WGRAD_ROWS;
    int nwg = 0;
    HF_HIP(hfn::launch_wgrad(L.cis, L.cos, w, &nwg, s));
    HF_HIP(hfn::launch_reduce_grad(n->part, nwg, L.cis, L.cos, L.ci, L.co, n->grads + L.off, n->part2, s));
    return HF_OK;
}

// Forward, loss and backward: grads and the loss (device) of one image.
int gradients(hf_net* n, const float* u, const float* target, int nx, int ny, hipStream_t s)
{
    const int64_t np = (int64_t)nx * ny;
    int rc = forward_hidden(n, u, np, nx, nx, ny, false, s);
    if (rc) return rc;
    hfn::ConvArgs a = conv_args(n, 4, n->act[3], nx, ny);
    a.outf = n->y;
    HF_HIP(hfn::launch_conv(32, 8, hfn::CONV_OUT, a, s));

    hfn::LossArgs la;
    std::memset(&la, 0, sizeof la);
    la.Y = n->y;
    la.Uc = u;
    la.T = target;
    for (int r = 0; r < 4; r++) la.inv_scale[r] = 1.0f / n->d.out_scale[r];
    const float inv = 1.0f / (4.0f * (float)np);
    la.gscale = inv;
    la.nx = nx;
    la.ny = ny;
    la.G = n->g8;
    la.part = n->lossp;
    int nb = 0;
    HF_HIP(hfn::launch_loss(la, &nb, s));
    HF_HIP(hfn::launch_reduce_sum(n->lossp, nb, inv, n->loss, s));

    // layer 5: gradient at its output is g8
    rc = wgrad(n, 4, n->act[3], n->g8, nx, ny, s);
    if (rc) return rc;
    const uint16_t* gin = n->g8;
    int cin = 8;
    for (int l = 4; l >= 1; l--) {
        // data gradient of layer l+1 (index l) into the pre-activation of layer l (index l-1)
        uint16_t* gout = n->gz[l & 1];
        hfn::ConvArgs b;
        std::memset(&b, 0, sizeof b);
        b.in = gin;
        b.wa = n->wb[l];
        b.nx = nx;
        b.ny = ny;
        b.rows = CONV_ROWS;
        b.out = gout;
        b.mask = n->act[l - 1];
        HF_HIP(hfn::launch_conv(cin, 32, hfn::CONV_BWD, b, s));
        const uint16_t* a_in = (l - 1 == 0) ? n->x : n->act[l - 2];
        rc = wgrad(n, l - 1, a_in, gout, nx, ny, s);
        if (rc) return rc;
        gin = gout;
        cin = 32;
    }
    return HF_OK;
}

}  // namespace

extern "C" {

int64_t hf_net_param_count(void) { return NPARAM; }

int hf_net_create(const hf_net_desc* desc, hf_net** out)
{
    if (!desc || !out) return fail(HF_EINVAL, "hf_net_create: null argument");
    *out = nullptr;
    hf_net* n = new hf_net();
    n->d = *desc;
    n->done_stream = (hipStream_t)-1;
    auto cleanup = [&](int rc) {
        hf_net_destroy(n);
        return rc;
    };
    if (int rc0 = hf_net_set_desc(n, desc)) return cleanup(rc0);
    if (hipMalloc(&n->params, NPARAM * sizeof(float)) != hipSuccess ||
        hipMalloc(&n->grads, NPARAM * sizeof(float)) != hipSuccess ||
        hipMalloc(&n->m, NPARAM * sizeof(float)) != hipSuccess ||
        hipMalloc(&n->v, NPARAM * sizeof(float)) != hipSuccess || hipMalloc(&n->loss, sizeof(float)) != hipSuccess ||
        hipMalloc(&n->fallback, sizeof(uint32_t)) != hipSuccess || hipEventCreate(&n->done) != hipSuccess)
        return cleanup(fail(HF_ENOMEM, "hf_net_create: device allocation failed"));
    for (int l = 0; l < NL; l++) {
        const Layer& L = LAYERS[l];
        if (hipMalloc(&n->wf[l], hfn::packed_weight_elems(L.cis, L.cos) * 2) != hipSuccess)
            return cleanup(fail(HF_ENOMEM, "hf_net_create: device allocation failed"));
        if (l > 0 && hipMalloc(&n->wb[l], hfn::packed_weight_elems(L.cos, L.cis) * 2) != hipSuccess)
            return cleanup(fail(HF_ENOMEM, "hf_net_create: device allocation failed"));
    }
    // He-normal weights (std sqrt(2 / (9 ci))), zero biases; the last layer
    // starts 10x smaller so the initial correction is small.
    std::vector<float> p(NPARAM, 0.0f);
    std::mt19937_64 rng(desc->seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    for (int l = 0; l < NL; l++) {
        const Layer& L = LAYERS[l];
        const double sd = std::sqrt(2.0 / (9.0 * L.ci)) * (l == NL - 1 ? 0.1 : 1.0);
        for (int64_t i = 0; i < 9 * L.ci * L.co; i++) p[L.off + i] = (float)(sd * nd(rng));
    }
    int rc = hf_net_set_params(n, p.data(), p.size());
    if (rc) return cleanup(rc);
    *out = n;
    return HF_OK;
}

void hf_net_destroy(hf_net* n)
{
    if (!n) return;
    if (hf_process_exiting()) {   // host memory only once the process is exiting
        delete n;
        return;
    }
    (void)hipDeviceSynchronize();
    free_images(n);
    (void)hipFree(n->params);
    (void)hipFree(n->grads);
    (void)hipFree(n->m);
    (void)hipFree(n->v);
    (void)hipFree(n->loss);
    (void)hipFree(n->fallback);
    for (int l = 0; l < NL; l++) {
        (void)hipFree(n->wf[l]);
        (void)hipFree(n->wb[l]);
    }
    if (n->done) (void)hipEventDestroy(n->done);
    delete n;
}

int hf_net_set_stream(hf_net* n, void* s)
{
    if (!n) return fail(HF_EINVAL, "null network");
    n->stream = (hipStream_t)s;
    return HF_OK;
}

int hf_net_set_desc(hf_net* n, const hf_net_desc* d)
{
    if (!n || !d) return fail(HF_EINVAL, "hf_net_set_desc: null argument");
    for (int r = 0; r < 4; r++)
        if (!(d->out_scale[r] != 0.0f) || !std::isfinite(d->out_scale[r]) || !std::isfinite(d->in_scale[r]) ||
            !std::isfinite(d->in_mean[r]))
            return fail(HF_EINVAL, "hf_net_set_desc: scales must be finite, out_scale non-zero");
    if (!std::isfinite(d->max_output) || !std::isfinite(d->shock_threshold) || !std::isfinite(d->gamma))
        return fail(HF_EINVAL, "hf_net_set_desc: gamma, shock_threshold and max_output must be finite");
    const uint64_t seed = n->d.seed;
    n->d = *d;
    n->d.seed = seed;
    return HF_OK;
}

int hf_net_set_params(hf_net* n, const float* p, size_t len)
{
    if (!n || !p) return fail(HF_EINVAL, "hf_net_set_params: null argument");
    if ((int64_t)len != NPARAM) return fail(HF_EINVAL, "hf_net_set_params: expected " + std::to_string(NPARAM) + " values");
    int rc = order_after(n, n->stream);
    if (rc) return rc;
    HF_HIP(hipMemcpyAsync(n->params, p, NPARAM * sizeof(float), hipMemcpyHostToDevice, n->stream));
    rc = pack_weights(n, n->stream);
    if (rc) return rc;
    rc = hf_net_reset_optimizer(n);
    if (rc) return rc;
    HF_HIP(hipStreamSynchronize(n->stream));
    return mark_done(n, n->stream);
}

int hf_net_get_params(hf_net* n, float* p, size_t len)
{
    if (!n || !p) return fail(HF_EINVAL, "hf_net_get_params: null argument");
    if ((int64_t)len != NPARAM) return fail(HF_EINVAL, "hf_net_get_params: expected " + std::to_string(NPARAM) + " values");
    int rc = order_after(n, n->stream);
    if (rc) return rc;
    HF_HIP(hipMemcpyAsync(p, n->params, NPARAM * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    HF_HIP(hipStreamSynchronize(n->stream));
    return mark_done(n, n->stream);
}

int hf_net_reset_optimizer(hf_net* n)
{
    if (!n) return fail(HF_EINVAL, "null network");
    HF_HIP(hipMemsetAsync(n->m, 0, NPARAM * sizeof(float), n->stream));
    HF_HIP(hipMemsetAsync(n->v, 0, NPARAM * sizeof(float), n->stream));
    n->step = 0;
    return HF_OK;
}

int hf_net_forward(hf_net* n, const float* u, int nx, int ny, float* y)
{
    if (!n || !u || !y) return fail(HF_EINVAL, "hf_net_forward: null argument");
    int rc = ensure_images(n, nx, ny);
    if (rc) return rc;
    rc = order_after(n, n->stream);
    if (rc) return rc;
    const int64_t np = (int64_t)nx * ny;
    rc = forward_hidden(n, u, np, nx, nx, ny, false, n->stream);
    if (rc) return rc;
    hfn::ConvArgs a = conv_args(n, 4, n->act[3], nx, ny);
    a.outf = y;
    HF_HIP(hfn::launch_conv(32, 8, hfn::CONV_OUT, a, n->stream));
    return mark_done(n, n->stream);
}

int hf_net_gradients(hf_net* n, const float* u, const float* target, int nx, int ny, float* grads, float* loss)
{
    if (!n || !u || !target) return fail(HF_EINVAL, "hf_net_gradients: null argument");
    int rc = ensure_images(n, nx, ny);
    if (rc) return rc;
    rc = order_after(n, n->stream);
    if (rc) return rc;
    rc = gradients(n, u, target, nx, ny, n->stream);
    if (rc) return rc;
    if (grads) HF_HIP(hipMemcpyAsync(grads, n->grads, NPARAM * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    if (loss) HF_HIP(hipMemcpyAsync(loss, n->loss, sizeof(float), hipMemcpyDeviceToHost, n->stream));
    HF_HIP(hipStreamSynchronize(n->stream));
    return mark_done(n, n->stream);
}

int hf_net_train_step(hf_net* n, const float* u, const float* target, int nx, int ny, float lr, float* loss_dev)
{
    if (!n || !u || !target) return fail(HF_EINVAL, "hf_net_train_step: null argument");
    if (!(lr > 0.0f) || !std::isfinite(lr)) return fail(HF_EINVAL, "hf_net_train_step: lr must be finite and > 0");
    int rc = ensure_images(n, nx, ny);
    if (rc) return rc;
    rc = order_after(n, n->stream);
    if (rc) return rc;
    rc = gradients(n, u, target, nx, ny, n->stream);
    if (rc) return rc;
    if (loss_dev) HF_HIP(hipMemcpyAsync(loss_dev, n->loss, sizeof(float), hipMemcpyDeviceToDevice, n->stream));
    n->step++;
    const double b1 = 0.9, b2 = 0.999;
    const float c1 = (float)(1.0 / (1.0 - std::pow(b1, (double)n->step)));
    const float c2 = (float)(1.0 / (1.0 - std::pow(b2, (double)n->step)));
    HF_HIP(hfn::launch_adam(n->params, n->grads, n->m, n->v, (int)NPARAM, lr, (float)b1, (float)b2, 1e-8f, c1, c2,
                            n->stream));
    rc = pack_weights(n, n->stream);
    if (rc) return rc;
    return mark_done(n, n->stream);
}

int hf_net_synchronize(hf_net* n)
{
    if (!n) return fail(HF_EINVAL, "null network");
    HF_HIP(hipStreamSynchronize(n->stream));
    return HF_OK;
}

// ------------------------------------------------------------ solver coupling --
static int fp32_block(hf_ctx* c, int b, const char* who)
{
    if (!c) return fail(HF_EINVAL, std::string(who) + ": null context");
    if (c->prec != HF_FP32) return fail(HF_EINVAL, std::string(who) + ": needs an FP32 context");
    if (b < 0 || b >= c->nb) return fail(HF_EINVAL, std::string(who) + ": block out of range");
    return HF_OK;
}

static float* interior(hf_ctx* c, int b)
{
    const HfBlockDev& B = c->blocks[b];
    return (float*)c->buf[c->cur] + B.off + (int64_t)B.g * B.pitch + B.g;
}

int hf_correct(hf_ctx* c, hf_net* n, int64_t* fallback)
{
    if (!n) return fail(HF_EINVAL, "hf_correct: null network");
    for (int b = 0; b < (c ? c->nb : 1); b++) {
        int rc = fp32_block(c, b, "hf_correct");
        if (rc) return rc;
        if (c->blocks[b].nx % 64 != 0) return fail(HF_EINVAL, "hf_correct: block nx must be a multiple of 64");
    }
    int rc = order_after(n, c->stream);
    if (rc) return rc;
    if (fallback) HF_HIP(hipMemsetAsync(n->fallback, 0, sizeof(uint32_t), c->stream));
    for (int b = 0; b < c->nb; b++) {
        const HfBlockDev& B = c->blocks[b];
        rc = ensure_images(n, B.nx, B.ny);
        if (rc) return rc;
        float* U = interior(c, b);
        const bool sensor = n->d.shock_threshold > 0.0f;
        rc = forward_hidden(n, U, B.plane, B.pitch, B.nx, B.ny, sensor, c->stream);
        if (rc) return rc;
        hfn::ConvArgs a = conv_args(n, 4, n->act[3], B.nx, B.ny);
        a.U = U;
        a.plane = B.plane;
        a.pitch = B.pitch;
        a.gamma = n->d.gamma;
        for (int r = 0; r < 4; r++) a.scale[r] = n->d.out_scale[r];
        a.max_output = n->d.max_output;
        a.shock = sensor ? n->shock : nullptr;
        a.fallback = n->fallback;
        HF_HIP(hfn::launch_conv(32, 8, hfn::CONV_APPLY, a, c->stream));
    }
    rc = mark_done(n, c->stream);
    if (rc) return rc;
    c->partials_valid = false;
    c->ghosts_pending = false;
    if (fallback) {
        uint32_t h = 0;
        HF_HIP(hipMemcpyAsync(&h, n->fallback, sizeof h, hipMemcpyDeviceToHost, c->stream));
        HF_HIP(hipStreamSynchronize(c->stream));
        *fallback = h;
    }
    return HF_OK;
}

int hf_state_to_device(hf_ctx* c, int b, float* dst)
{
    int rc = fp32_block(c, b, "hf_state_to_device");
    if (rc) return rc;
    if (!dst) return fail(HF_EINVAL, "hf_state_to_device: null destination");
    const HfBlockDev& B = c->blocks[b];
    HF_HIP(hfn::launch_state_copy(interior(c, b), B.plane, B.pitch, dst, (int64_t)B.nx * B.ny, B.nx, B.nx, B.ny,
                                  c->stream));
    return HF_OK;
}

int hf_state_from_device(hf_ctx* c, int b, const float* src)
{
    int rc = fp32_block(c, b, "hf_state_from_device");
    if (rc) return rc;
    if (!src) return fail(HF_EINVAL, "hf_state_from_device: null source");
    const HfBlockDev& B = c->blocks[b];
    HF_HIP(hfn::launch_state_copy(src, (int64_t)B.nx * B.ny, B.nx, interior(c, b), B.plane, B.pitch, B.nx, B.ny,
                                  c->stream));
    c->partials_valid = false;
    c->ghosts_pending = false;
    return HF_OK;
}

int hf_restrict2(hf_ctx* c, int b, float* dst)
{
    int rc = fp32_block(c, b, "hf_restrict2");
    if (rc) return rc;
    if (!dst) return fail(HF_EINVAL, "hf_restrict2: null destination");
    const HfBlockDev& B = c->blocks[b];
    if (B.nx % 2 || B.ny % 2) return fail(HF_EINVAL, "hf_restrict2: block sides must be even");
    HF_HIP(hfn::launch_restrict2(interior(c, b), B.plane, B.pitch, B.nx / 2, B.ny / 2, dst, c->stream));
    return HF_OK;
}

void* hf_stream(hf_ctx* c) { return c ? (void*)c->stream : nullptr; }

}  // extern "C"
