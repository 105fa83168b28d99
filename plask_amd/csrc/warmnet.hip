// Learned warm start: a 3-level 3D convolutional encoder-decoder (channels
// 32/64/128, 3x3x3 kernels, bf16 weights and activations, fp32 accumulation on
// MFMA) mapping per-node input fields of a layered stack to a normalised
// temperature guess, trained on the GPU with the hand-written kernels of
// warmnet_kernels.hpp.  No counterpart in the reference.
//
//   e1 = relu(C1(x) + b1)            level 1 (nx, ny, nz), 32 ch
//   e2 = relu(C2(pool(e1)) + b2)     level 2, 64 ch
//   e3 = relu(C3(pool(e2)) + b3)     level 3, 128 ch
//   d2 = relu(C4(up(e3)) + b4 + e2)  level 2, 64 ch
//   d1 = relu(C5(up(d2)) + b5 + e1)  level 1, 32 ch
//   y  = h . d1 + hb                 level 1, fp32
// Parameters (fp32 master, flat): per conv w[27][cout][cin] then b[cout]; then
// h[32], hb.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/plask_warmnet.h"
#include "warmnet_kernels.hpp"

using namespace wn;

namespace {
struct Layer {
    int cin, cout, level;
    size_t w_off, b_off;
    bf16 *wf = nullptr, *wt = nullptr;
};
constexpr int NIN = 8;  // input channels
}  // namespace

struct pf_net {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    Vol L[3];
    Layer layer[5];
    size_t h_off = 0, hb_off = 0, nparams = 0;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    int adam_t = 0;
    // activations
    bf16 *x = nullptr, *e1 = nullptr, *p1 = nullptr, *e2 = nullptr, *p2 = nullptr, *e3 = nullptr, *d2 = nullptr,
         *d1 = nullptr;
    float* y = nullptr;
    float* gy = nullptr;   // dL/dy
    float lambda = 0.f;    // weight of the neighbour-difference term of the loss
    // gradients of activations
    bf16 *g5 = nullptr, *gu1 = nullptr, *g4 = nullptr, *gu2 = nullptr, *g3 = nullptr, *gp2 = nullptr, *gp1 = nullptr;
    float* fin = nullptr;  // fp32 staging of host inputs
    // dataset
    int max_samples = 0;
    bf16* dfeat = nullptr;
    float* dtarget = nullptr;
    float* dwz = nullptr;            // per-sample z-edge loss weights (optional)
    std::vector<char> has_wz;
    float* mask = nullptr;
    float mask_norm = 1.f;
    double* dloss = nullptr;
    bool weights_ready = false;
    std::vector<void*> allocs;
};

#define NCK(call)                                                                \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess) {                                                  \
            net->err = std::string(#call) + ": " + hipGetErrorString(e_);        \
            return PF_NET_EDEVICE;                                               \
        }                                                                        \
    } while (0)

static int nfail(pf_net* net, int code, const std::string& msg) {
    net->err = msg;
    return code;
}

template <typename T>
static int nalloc(pf_net* net, T** ptr, size_t count) {
    NCK(hipMalloc((void**)ptr, sizeof(T) * std::max<size_t>(count, 1)));
    net->allocs.push_back((void*)*ptr);
    return PF_NET_OK;
}

static int gridn(long long n, int threads = 256, int cap = 8192) {
    return (int)std::max<long long>(1, std::min<long long>(cap, (n + threads - 1) / threads));
}

// ---------------------------------------------------------------------------
// convolution dispatch (CIN and the channel-group width are template arguments)
// ---------------------------------------------------------------------------
template <int CIN, bool UP, bool RELU>
static void conv_cin(pf_net* net, const Vol& o, const bf16* in, const bf16* w, const float* bias, const bf16* skip,
                     bf16* out, int cout) {
    const long long tiles = o.n() / 16;
    if (cout % 64 == 0) {
        const long long waves = tiles * (cout / 64);
        hipLaunchKernelGGL((k_conv3<CIN, 4, UP, RELU>), dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, net->stream, o,
                           in, w, bias, skip, out, cout);
    } else {
        const long long waves = tiles * (cout / 32);
        hipLaunchKernelGGL((k_conv3<CIN, 2, UP, RELU>), dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, net->stream, o,
                           in, w, bias, skip, out, cout);
    }
}

template <bool UP, bool RELU>
static void conv(pf_net* net, const Vol& o, int cin, const bf16* in, const bf16* w, const float* bias, const bf16* skip,
                 bf16* out, int cout) {
    switch (cin) {
        case 8: conv_cin<8, UP, RELU>(net, o, in, w, bias, skip, out, cout); break;
        case 32: conv_cin<32, UP, RELU>(net, o, in, w, bias, skip, out, cout); break;
        case 64: conv_cin<64, UP, RELU>(net, o, in, w, bias, skip, out, cout); break;
        case 128: conv_cin<128, UP, RELU>(net, o, in, w, bias, skip, out, cout); break;
    }
}

template <bool UP>
static void wgrad(pf_net* net, const Vol& o, const bf16* gy, const bf16* x, float* gw, int cout, int cin) {
    const int vchunk = 4096;
    const long long nch = (o.n() + vchunk - 1) / vchunk;
    const long long waves = 27LL * (cout / 16) * ((cin + 15) / 16) * nch;
    hipLaunchKernelGGL((k_conv3_wgrad<UP>), dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, net->stream, o, gy, x, gw,
                       cout, cin, vchunk);
}

static void chan_sum(pf_net* net, long long nv, int c, const bf16* g, float* gb) {
    hipLaunchKernelGGL(k_chan_sum, dim3(gridn(nv * c, 256, 2048)), dim3(256), 0, net->stream, nv, c, g, gb);
}

static void prep_weights(pf_net* net) {
    for (int l = 0; l < 5; ++l) {
        Layer& L = net->layer[l];
        const long long n = 27LL * L.cout * L.cin;
        hipLaunchKernelGGL(k_prep_weights, dim3(gridn(n)), dim3(256), 0, net->stream, net->p + L.w_off, L.cout, L.cin,
                           L.cin, L.wf, L.wt);
    }
    net->weights_ready = true;
}

// forward pass on net->x
static void forward(pf_net* net) {
    if (!net->weights_ready) prep_weights(net);
    const Vol &V1 = net->L[0], &V2 = net->L[1], &V3 = net->L[2];
    Layer* Ly = net->layer;
    float* P = net->p;
    conv<false, true>(net, V1, NIN, net->x, Ly[0].wf, P + Ly[0].b_off, nullptr, net->e1, 32);
    hipLaunchKernelGGL(k_pool, dim3(gridn(V2.n() * 32)), dim3(256), 0, net->stream, V2, net->e1, net->p1, 32);
    conv<false, true>(net, V2, 32, net->p1, Ly[1].wf, P + Ly[1].b_off, nullptr, net->e2, 64);
    hipLaunchKernelGGL(k_pool, dim3(gridn(V3.n() * 64)), dim3(256), 0, net->stream, V3, net->e2, net->p2, 64);
    conv<false, true>(net, V3, 64, net->p2, Ly[2].wf, P + Ly[2].b_off, nullptr, net->e3, 128);
    conv<true, true>(net, V2, 128, net->e3, Ly[3].wf, P + Ly[3].b_off, net->e2, net->d2, 64);
    conv<true, true>(net, V1, 64, net->d2, Ly[4].wf, P + Ly[4].b_off, net->e1, net->d1, 32);
    hipLaunchKernelGGL(k_head, dim3(gridn(V1.n())), dim3(256), 0, net->stream, V1.n(), net->d1, P + net->h_off,
                       P + net->hb_off, net->y);
}

// backward pass against target t (fp32 per node); accumulates into net->g
static void backward(pf_net* net, const float* t, const float* wz) {
    const Vol &V1 = net->L[0], &V2 = net->L[1], &V3 = net->L[2];
    Layer* Ly = net->layer;
    float* P = net->p;
    float* G = net->g;
    hipLaunchKernelGGL(k_loss, dim3(gridn(V1.n(), 256, 2048)), dim3(256), 0, net->stream, V1, net->y, t, net->mask, wz,
                       net->mask_norm, net->lambda, net->gy, net->dloss);
    hipLaunchKernelGGL(k_head_bwd, dim3(gridn(V1.n(), 256, 1024)), dim3(256), 0, net->stream, V1.n(), net->gy,
                       net->d1, P + net->h_off, net->g5, G + net->h_off, G + net->hb_off);
    // C5: d1 = relu(C5(up(d2)) + b5 + e1); g5 = gradient of its pre-activation
    chan_sum(net, V1.n(), 32, net->g5, G + Ly[4].b_off);
    wgrad<true>(net, V1, net->g5, net->d2, G + Ly[4].w_off, 32, 64);
    conv<false, false>(net, V1, 32, net->g5, Ly[4].wt, nullptr, nullptr, net->gu1, 64);
    hipLaunchKernelGGL(k_up_bwd, dim3(gridn(V2.n() * 64)), dim3(256), 0, net->stream, V2, net->gu1, net->g4, 64,
                       net->d2);
    // C4: d2 = relu(C4(up(e3)) + b4 + e2)
    chan_sum(net, V2.n(), 64, net->g4, G + Ly[3].b_off);
    wgrad<true>(net, V2, net->g4, net->e3, G + Ly[3].w_off, 64, 128);
    conv<false, false>(net, V2, 64, net->g4, Ly[3].wt, nullptr, nullptr, net->gu2, 128);
    hipLaunchKernelGGL(k_up_bwd, dim3(gridn(V3.n() * 128)), dim3(256), 0, net->stream, V3, net->gu2, net->g3, 128,
                       net->e3);
    // C3: e3 = relu(C3(p2) + b3)
    chan_sum(net, V3.n(), 128, net->g3, G + Ly[2].b_off);
    wgrad<false>(net, V3, net->g3, net->p2, G + Ly[2].w_off, 128, 64);
    conv<false, false>(net, V3, 128, net->g3, Ly[2].wt, nullptr, nullptr, net->gp2, 64);
    // e2 feeds the skip of d2 (gradient g4) and pool -> p2 (gradient gp2); in place into g4
    hipLaunchKernelGGL(k_skip_pool_bwd, dim3(gridn(V2.n() * 64)), dim3(256), 0, net->stream, V2, net->g4, net->gp2,
                       net->e2, net->g4, 64);
    // C2: e2 = relu(C2(p1) + b2)
    chan_sum(net, V2.n(), 64, net->g4, G + Ly[1].b_off);
    wgrad<false>(net, V2, net->g4, net->p1, G + Ly[1].w_off, 64, 32);
    conv<false, false>(net, V2, 64, net->g4, Ly[1].wt, nullptr, nullptr, net->gp1, 32);
    hipLaunchKernelGGL(k_skip_pool_bwd, dim3(gridn(V1.n() * 32)), dim3(256), 0, net->stream, V1, net->g5, net->gp1,
                       net->e1, net->g5, 32);
    // C1: e1 = relu(C1(x) + b1)
    chan_sum(net, V1.n(), 32, net->g5, G + Ly[0].b_off);
    wgrad<false>(net, V1, net->g5, net->x, G + Ly[0].w_off, 32, NIN);
}

static int upload_feat(pf_net* net, const float* feat) {
    const long long n = net->L[0].n() * NIN;
    NCK(hipMemcpyAsync(net->fin, feat, sizeof(float) * n, hipMemcpyHostToDevice, net->stream));
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(gridn(n)), dim3(256), 0, net->stream, n, net->fin, net->x);
    return PF_NET_OK;
}

static const float* sample_wz(pf_net* net, int idx) {
    return (net->dwz && net->has_wz[idx]) ? net->dwz + (size_t)idx * net->L[0].n() : nullptr;
}

static void load_sample(pf_net* net, int idx) {
    const long long n = net->L[0].n() * NIN;
    hipMemcpyAsync(net->x, net->dfeat + (size_t)idx * n, sizeof(bf16) * n, hipMemcpyDeviceToDevice, net->stream);
}

extern "C" {

int pf_net_create(pf_net** out, int device, int nx, int ny, int nz, int max_samples) {
    if (!out) return PF_NET_EINVAL;
    *out = nullptr;
    if (nx % 4 || ny % 4 || nz % 4 || nx < 4 || ny < 4 || nz < 4 || max_samples < 0) return PF_NET_EINVAL;
    const long long v3 = (long long)(nx / 4) * (ny / 4) * (nz / 4);
    if (v3 % 32) return PF_NET_EINVAL;  // every level a whole number of 32-voxel steps
    pf_net* net = new pf_net();
    net->device = device;
    auto bail = [&](int rc) {
        pf_net_destroy(net);
        return rc;
    };
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&net->stream) != hipSuccess) return bail(PF_NET_EDEVICE);
    net->L[0] = Vol{nx, ny, nz};
    net->L[1] = Vol{nx / 2, ny / 2, nz / 2};
    net->L[2] = Vol{nx / 4, ny / 4, nz / 4};
    const int spec[5][3] = {{NIN, 32, 0}, {32, 64, 1}, {64, 128, 2}, {128, 64, 1}, {64, 32, 0}};
    size_t off = 0;
    for (int l = 0; l < 5; ++l) {
        Layer& L = net->layer[l];
        L.cin = spec[l][0];
        L.cout = spec[l][1];
        L.level = spec[l][2];
        L.w_off = off;
        off += 27 * (size_t)L.cin * L.cout;
        L.b_off = off;
        off += L.cout;
    }
    net->h_off = off;
    off += 32;
    net->hb_off = off;
    off += 1;
    net->nparams = off;
    int rc;
    for (float** pp : {&net->p, &net->g, &net->m, &net->v})
        if ((rc = nalloc(net, pp, net->nparams))) return bail(rc);
    for (int l = 0; l < 5; ++l) {
        Layer& L = net->layer[l];
        if ((rc = nalloc(net, &L.wf, 27 * (size_t)L.cin * L.cout))) return bail(rc);
        if (l > 0 && (rc = nalloc(net, &L.wt, 27 * (size_t)L.cin * L.cout))) return bail(rc);
    }
    const size_t n1 = net->L[0].n(), n2 = net->L[1].n(), n3 = net->L[2].n();
    struct {
        bf16** p;
        size_t n;
    } acts[] = {{&net->x, n1 * NIN}, {&net->e1, n1 * 32}, {&net->p1, n2 * 32}, {&net->e2, n2 * 64},
                {&net->p2, n3 * 64}, {&net->e3, n3 * 128}, {&net->d2, n2 * 64}, {&net->d1, n1 * 32},
                {&net->g5, n1 * 32}, {&net->gu1, n1 * 64}, {&net->g4, n2 * 64}, {&net->gu2, n2 * 128},
                {&net->g3, n3 * 128}, {&net->gp2, n3 * 64}, {&net->gp1, n2 * 32}};
    for (auto& a : acts)
        if ((rc = nalloc(net, a.p, a.n))) return bail(rc);
    if ((rc = nalloc(net, &net->y, n1))) return bail(rc);
    if ((rc = nalloc(net, &net->gy, n1))) return bail(rc);
    if ((rc = nalloc(net, &net->fin, n1 * NIN))) return bail(rc);
    if ((rc = nalloc(net, &net->mask, n1))) return bail(rc);
    if ((rc = nalloc(net, &net->dloss, 1))) return bail(rc);
    net->max_samples = max_samples;
    if (max_samples) {
        if ((rc = nalloc(net, &net->dfeat, n1 * NIN * (size_t)max_samples))) return bail(rc);
        if ((rc = nalloc(net, &net->dtarget, n1 * (size_t)max_samples))) return bail(rc);
    }
    // mask = 1 everywhere until set
    std::vector<float> ones(n1, 1.f);
    if (hipMemcpy(net->mask, ones.data(), sizeof(float) * n1, hipMemcpyHostToDevice) != hipSuccess)
        return bail(PF_NET_EDEVICE);
    net->mask_norm = 1.f / (float)n1;
    for (float* q : {net->g, net->m, net->v})
        hipLaunchKernelGGL(k_zero_f32, dim3(gridn(net->nparams)), dim3(256), 0, net->stream, (long long)net->nparams, q);
    if (hipStreamSynchronize(net->stream) != hipSuccess) return bail(PF_NET_EDEVICE);
    *out = net;
    return PF_NET_OK;
}

void pf_net_destroy(pf_net* net) {
    if (!net) return;
    hipSetDevice(net->device);
    if (net->stream) hipStreamSynchronize(net->stream);
    for (void* a : net->allocs) hipFree(a);
    if (net->stream) hipStreamDestroy(net->stream);
    delete net;
}

const char* pf_net_last_error(const pf_net* net) { return net ? net->err.c_str() : "null network"; }

int64_t pf_net_num_params(const pf_net* net) { return net ? (int64_t)net->nparams : -1; }

int pf_net_set_params(pf_net* net, const float* params, int64_t n) {
    if (!net || !params || n != (int64_t)net->nparams) return PF_NET_EINVAL;
    NCK(hipSetDevice(net->device));
    NCK(hipMemcpy(net->p, params, sizeof(float) * n, hipMemcpyHostToDevice));
    for (float* q : {net->m, net->v})
        hipLaunchKernelGGL(k_zero_f32, dim3(gridn(n)), dim3(256), 0, net->stream, (long long)n, q);
    net->adam_t = 0;
    net->weights_ready = false;
    NCK(hipStreamSynchronize(net->stream));
    return PF_NET_OK;
}

int pf_net_get_params(pf_net* net, float* params, int64_t n) {
    if (!net || !params || n != (int64_t)net->nparams) return PF_NET_EINVAL;
    NCK(hipSetDevice(net->device));
    NCK(hipStreamSynchronize(net->stream));
    NCK(hipMemcpy(params, net->p, sizeof(float) * n, hipMemcpyDeviceToHost));
    return PF_NET_OK;
}

int pf_net_set_mask(pf_net* net, const float* mask) {
    if (!net || !mask) return PF_NET_EINVAL;
    NCK(hipSetDevice(net->device));
    const size_t n1 = net->L[0].n();
    double s = 0.;
    for (size_t i = 0; i < n1; ++i) s += mask[i];
    if (!(s > 0.)) return nfail(net, PF_NET_EINVAL, "mask selects no node");
    NCK(hipMemcpy(net->mask, mask, sizeof(float) * n1, hipMemcpyHostToDevice));
    net->mask_norm = (float)(1. / s);
    return PF_NET_OK;
}

int pf_net_sample_weights(pf_net* net, int idx, const float* wz) {
    if (!net || !wz) return PF_NET_EINVAL;
    if (idx < 0 || idx >= net->max_samples) return nfail(net, PF_NET_EINVAL, "sample index out of range");
    NCK(hipSetDevice(net->device));
    const size_t n1 = net->L[0].n();
    if (!net->dwz) {
        int rc;
        if ((rc = nalloc(net, &net->dwz, n1 * (size_t)net->max_samples))) return rc;
        net->has_wz.assign(net->max_samples, 0);
    }
    NCK(hipMemcpy(net->dwz + (size_t)idx * n1, wz, sizeof(float) * n1, hipMemcpyHostToDevice));
    net->has_wz[idx] = 1;
    return PF_NET_OK;
}

int pf_net_set_loss(pf_net* net, float lambda) {
    if (!net || !(lambda >= 0.f)) return PF_NET_EINVAL;
    net->lambda = lambda;
    return PF_NET_OK;
}

int pf_net_sample_set(pf_net* net, int idx, const float* feat, const float* target) {
    if (!net || !feat || !target) return PF_NET_EINVAL;
    if (idx < 0 || idx >= net->max_samples) return nfail(net, PF_NET_EINVAL, "sample index out of range");
    NCK(hipSetDevice(net->device));
    int rc;
    if ((rc = upload_feat(net, feat))) return rc;
    const long long n1 = net->L[0].n();
    NCK(hipMemcpyAsync(net->dfeat + (size_t)idx * n1 * NIN, net->x, sizeof(bf16) * n1 * NIN, hipMemcpyDeviceToDevice,
                       net->stream));
    NCK(hipMemcpyAsync(net->dtarget + (size_t)idx * n1, target, sizeof(float) * n1, hipMemcpyHostToDevice, net->stream));
    NCK(hipStreamSynchronize(net->stream));
    return PF_NET_OK;
}

int pf_net_predict(pf_net* net, const float* feat, float* out) {
    if (!net || !feat) return PF_NET_EINVAL;
    NCK(hipSetDevice(net->device));
    int rc;
    if ((rc = upload_feat(net, feat))) return rc;
    forward(net);
    NCK(hipGetLastError());
    if (out) NCK(hipMemcpyAsync(out, net->y, sizeof(float) * net->L[0].n(), hipMemcpyDeviceToHost, net->stream));
    NCK(hipStreamSynchronize(net->stream));
    return PF_NET_OK;
}

const float* pf_net_output_device(const pf_net* net) { return net ? net->y : nullptr; }

int pf_net_predict_sample(pf_net* net, int idx, float* out) {
    if (!net || !out) return PF_NET_EINVAL;
    if (idx < 0 || idx >= net->max_samples) return nfail(net, PF_NET_EINVAL, "sample index out of range");
    NCK(hipSetDevice(net->device));
    load_sample(net, idx);
    forward(net);
    NCK(hipGetLastError());
    NCK(hipMemcpyAsync(out, net->y, sizeof(float) * net->L[0].n(), hipMemcpyDeviceToHost, net->stream));
    NCK(hipStreamSynchronize(net->stream));
    return PF_NET_OK;
}

int pf_net_gradient(pf_net* net, int idx, float* grads, double* loss) {
    if (!net || !grads) return PF_NET_EINVAL;
    if (idx < 0 || idx >= net->max_samples) return nfail(net, PF_NET_EINVAL, "sample index out of range");
    NCK(hipSetDevice(net->device));
    hipLaunchKernelGGL(k_zero_f32, dim3(gridn(net->nparams)), dim3(256), 0, net->stream, (long long)net->nparams, net->g);
    hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, net->stream, 1LL, net->dloss);
    load_sample(net, idx);
    forward(net);
    backward(net, net->dtarget + (size_t)idx * net->L[0].n(), sample_wz(net, idx));
    NCK(hipGetLastError());
    NCK(hipMemcpyAsync(grads, net->g, sizeof(float) * net->nparams, hipMemcpyDeviceToHost, net->stream));
    double l = 0.;
    NCK(hipMemcpyAsync(&l, net->dloss, sizeof(double), hipMemcpyDeviceToHost, net->stream));
    NCK(hipStreamSynchronize(net->stream));
    if (loss) *loss = l;
    return PF_NET_OK;
}

int pf_net_train_batch(pf_net* net, const int* idx, int n, float lr, double* loss) {
    if (!net || !idx || n <= 0) return PF_NET_EINVAL;
    for (int i = 0; i < n; ++i)
        if (idx[i] < 0 || idx[i] >= net->max_samples) return nfail(net, PF_NET_EINVAL, "sample index out of range");
    NCK(hipSetDevice(net->device));
    hipLaunchKernelGGL(k_zero_f32, dim3(gridn(net->nparams)), dim3(256), 0, net->stream, (long long)net->nparams, net->g);
    hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, net->stream, 1LL, net->dloss);
    for (int i = 0; i < n; ++i) {
        load_sample(net, idx[i]);
        forward(net);
        backward(net, net->dtarget + (size_t)idx[i] * net->L[0].n(), sample_wz(net, idx[i]));
    }
    net->adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f / (1.f - std::pow(b1, (float)net->adam_t));
    const float c2 = 1.f / (1.f - std::pow(b2, (float)net->adam_t));
    hipLaunchKernelGGL(k_adam, dim3(gridn(net->nparams)), dim3(256), 0, net->stream, (long long)net->nparams, net->p,
                       net->g, net->m, net->v, lr, b1, b2, 1e-8f, c1, c2, 1.f / (float)n);
    prep_weights(net);
    NCK(hipGetLastError());
    double l = 0.;
    NCK(hipMemcpyAsync(&l, net->dloss, sizeof(double), hipMemcpyDeviceToHost, net->stream));
    NCK(hipStreamSynchronize(net->stream));
    if (loss) *loss = l / n;
    return PF_NET_OK;
}

}  // extern "C"
