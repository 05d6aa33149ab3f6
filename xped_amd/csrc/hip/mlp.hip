// Small fully connected network in FP32 for the learned rank predictor: nin -> width -> width -> 1
// with ReLU after the hidden layers, mean-squared-error loss, Adam.  Forward, backward and the
// optimiser step are plain per-element kernels: the network is tiny (a few thousand weights) and
// the batch (sector samples) is at most some ten thousand rows, so every launch is latency bound.
//
// Parameter layout (one FP32 array): W1 (nin x width, row i = input i), b1 (width),
// W2 (width x width), b2 (width), w3 (width), b3 (1).
#include <cmath>

#include "device.hpp"

using xh::check;
using xh::guard;

struct xh_mlp {
    int nin = 0, width = 0;
    int64_t nparams = 0;
    float* params = nullptr;  // device
    float* m = nullptr;       // Adam first moment
    float* v = nullptr;       // Adam second moment
    int64_t cap = 0;          // batch capacity of the activation buffers
    float* h1 = nullptr;      // B x width (post-activation)
    float* h2 = nullptr;      // B x width
    float* d1 = nullptr;      // B x width (gradients w.r.t. pre-activations)
    float* d2 = nullptr;
    float* dy = nullptr;      // B
    float* y = nullptr;       // B
    float* loss_part = nullptr;
};

namespace {

constexpr int MT = 256;

struct Layout {
    int64_t W1, b1, W2, b2, w3, b3, n;
};
Layout layout(int nin, int w)
{
    Layout L{};
    L.W1 = 0;
    L.b1 = L.W1 + int64_t(nin) * w;
    L.W2 = L.b1 + w;
    L.b2 = L.W2 + int64_t(w) * w;
    L.w3 = L.b2 + w;
    L.b3 = L.w3 + w;
    L.n = L.b3 + 1;
    return L;
}

// out[b, j] = relu(sum_i in[b, i] W[i, j] + bias[j]); in is B x nin row-major
__global__ void k_dense_relu(const float* __restrict__ in, int nin, const float* __restrict__ W, const float* __restrict__ bias, int nout,
                             int64_t B, float* __restrict__ out)
{
    const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(e >= B * nout) return;
    const int64_t b = e / nout;
    const int j = static_cast<int>(e % nout);
    float acc = bias[j];
    for(int i = 0; i < nin; ++i) acc = fmaf(in[b * nin + i], W[int64_t(i) * nout + j], acc);
    out[e] = acc > 0.f ? acc : 0.f;
}

// y[b] = sum_j h[b, j] w[j] + c
__global__ void k_dense_out(const float* __restrict__ h, int w, const float* __restrict__ wv, const float* __restrict__ c, int64_t B, float* __restrict__ y)
{
    const int64_t b = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(b >= B) return;
    float acc = c[0];
    for(int j = 0; j < w; ++j) acc = fmaf(h[b * w + j], wv[j], acc);
    y[b] = acc;
}

// dy = 2 (y - t) / B and per-block partial sums of (y - t)^2
__global__ void k_mse(const float* __restrict__ y, const float* __restrict__ t, int64_t B, float* __restrict__ dy, float* __restrict__ part)
{
    __shared__ float s[MT];
    const int64_t b = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    float r = 0.f;
    if(b < B) {
        const float d = y[b] - t[b];
        dy[b] = 2.f * d / static_cast<float>(B);
        r = d * d;
    }
    s[threadIdx.x] = r;
    __syncthreads();
    for(int o = MT / 2; o > 0; o >>= 1) {
        if(threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if(threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// gradient of the output layer: gw3[j] = sum_b h2[b, j] dy[b], gb3 = sum_b dy[b], and
// d2[b, j] = dy[b] w3[j] masked by the ReLU of layer 2
__global__ void k_back_out(const float* __restrict__ h2, const float* __restrict__ dy, const float* __restrict__ w3, int w, int64_t B, float* __restrict__ gw3,
                           float* __restrict__ gb3, float* __restrict__ d2)
{
    const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(e < B * w) {
        const int64_t b = e / w;
        const int j = static_cast<int>(e % w);
        d2[e] = h2[e] > 0.f ? dy[b] * w3[j] : 0.f;
    }
    if(e < w) {
        float acc = 0.f;
        for(int64_t b = 0; b < B; ++b) acc = fmaf(h2[b * w + e], dy[b], acc);
        gw3[e] = acc;
    }
    if(e == 0) {
        float acc = 0.f;
        for(int64_t b = 0; b < B; ++b) acc += dy[b];
        gb3[0] = acc;
    }
}

// weight gradient of a dense layer: gW[i, j] = sum_b in[b, i] d[b, j], gb[j] = sum_b d[b, j]
__global__ void k_back_weights(const float* __restrict__ in, int nin, const float* __restrict__ d, int nout, int64_t B, float* __restrict__ gW,
                               float* __restrict__ gb)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if(e < nin * nout) {
        const int i = e / nout, j = e % nout;
        float acc = 0.f;
        for(int64_t b = 0; b < B; ++b) acc = fmaf(in[b * nin + i], d[b * nout + j], acc);
        gW[e] = acc;
    }
    if(e < nout) {
        float acc = 0.f;
        for(int64_t b = 0; b < B; ++b) acc += d[b * nout + e];
        gb[e] = acc;
    }
}

// d_prev[b, i] = (sum_j d[b, j] W[i, j]) masked by the ReLU of the previous layer's output h
__global__ void k_back_input(const float* __restrict__ d, const float* __restrict__ W, int nin, int nout, const float* __restrict__ h, int64_t B,
                             float* __restrict__ dprev)
{
    const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(e >= B * nin) return;
    const int64_t b = e / nin;
    const int i = static_cast<int>(e % nin);
    if(!(h[e] > 0.f)) {
        dprev[e] = 0.f;
        return;
    }
    float acc = 0.f;
    for(int j = 0; j < nout; ++j) acc = fmaf(d[b * nout + j], W[int64_t(i) * nout + j], acc);
    dprev[e] = acc;
}

// Adam (Kingma & Ba, ICLR 2015) with bias correction
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v, int64_t n, float lr, float b1,
                       float b2, float eps, float c1, float c2)
{
    const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if(e >= n) return;
    const float gi = g[e];
    const float mi = b1 * m[e] + (1.f - b1) * gi;
    const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    p[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + MT - 1) / MT); }

void ensure_capacity(xh_ctx* ctx, xh_mlp* net, int64_t B)
{
    if(B <= net->cap) return;
    for(float* p : {net->h1, net->h2, net->d1, net->d2, net->dy, net->y, net->loss_part}) xh::dev_free(ctx, p);
    const int64_t w = net->width;
    auto a = [&](int64_t n) {
        float* p = nullptr;
        check(xh::dmalloc(ctx, reinterpret_cast<void**>(&p), size_t(std::max<int64_t>(n, 1)) * sizeof(float)), "hipMallocAsync(mlp)");
        return p;
    };
    net->h1 = a(B * w);
    net->h2 = a(B * w);
    net->d1 = a(B * w);
    net->d2 = a(B * w);
    net->dy = a(B);
    net->y = a(B);
    net->loss_part = a(blocks_for(B));
    net->cap = B;
}

void forward(xh_ctx* ctx, xh_mlp* net, const float* X, int64_t B, float* y)
{
    const Layout L = layout(net->nin, net->width);
    const int w = net->width;
    hipLaunchKernelGGL(k_dense_relu, dim3(blocks_for(B * w)), dim3(MT), 0, ctx->stream, X, net->nin, net->params + L.W1, net->params + L.b1, w, B, net->h1);
    hipLaunchKernelGGL(k_dense_relu, dim3(blocks_for(B * w)), dim3(MT), 0, ctx->stream, net->h1, w, net->params + L.W2, net->params + L.b2, w, B, net->h2);
    hipLaunchKernelGGL(k_dense_out, dim3(blocks_for(B)), dim3(MT), 0, ctx->stream, net->h2, w, net->params + L.w3, net->params + L.b3, B, y);
    xh::launch_check("mlp forward");
}

}  // namespace

extern "C" int xh_mlp_create(xh_ctx* ctx, int nin, int width, uint64_t seed, xh_mlp** out)
{
    *out = nullptr;
    auto* net = new xh_mlp();
    const int rc = guard(ctx, [&] {
        if(nin <= 0 || width <= 0 || width > 1024) throw xh::Error(XH_EINVAL, "mlp: invalid shape");
        net->nin = nin;
        net->width = width;
        const Layout L = layout(nin, width);
        net->nparams = L.n;
        // He-normal weights from a splitmix64 + Box-Muller stream, zero biases
        std::vector<float> h(L.n, 0.f);
        uint64_t st = seed;
        auto next = [&]() {
            uint64_t z = (st += 0x9e3779b97f4a7c15ull);
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            return z ^ (z >> 31);
        };
        auto normal = [&]() {
            const double u1 = (double((next() >> 11) + 1)) * 0x1.0p-53, u2 = double(next() >> 11) * 0x1.0p-53;
            return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
        };
        for(int64_t i = 0; i < int64_t(nin) * width; ++i) h[L.W1 + i] = float(normal() * std::sqrt(2.0 / nin));
        for(int64_t i = 0; i < int64_t(width) * width; ++i) h[L.W2 + i] = float(normal() * std::sqrt(2.0 / width));
        for(int64_t i = 0; i < width; ++i) h[L.w3 + i] = float(normal() * std::sqrt(1.0 / width));
        for(float** p : {&net->params, &net->m, &net->v})
            check(xh::dmalloc(ctx, reinterpret_cast<void**>(p), size_t(L.n) * sizeof(float)), "hipMallocAsync(mlp)");
        check(hipMemcpyAsync(net->params, h.data(), size_t(L.n) * sizeof(float), hipMemcpyHostToDevice, ctx->stream), "hipMemcpyAsync");
        check(hipMemsetAsync(net->m, 0, size_t(L.n) * sizeof(float), ctx->stream), "hipMemsetAsync");
        check(hipMemsetAsync(net->v, 0, size_t(L.n) * sizeof(float), ctx->stream), "hipMemsetAsync");
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
    if(rc != XH_OK) {
        delete net;
        return rc;
    }
    *out = net;
    return XH_OK;
}

extern "C" int xh_mlp_destroy(xh_ctx* ctx, xh_mlp* net)
{
    if(!net) return XH_OK;
    return guard(ctx, [&] {
        for(float* p : {net->params, net->m, net->v, net->h1, net->h2, net->d1, net->d2, net->dy, net->y, net->loss_part}) xh::dev_free(ctx, p);
        delete net;
    });
}

extern "C" int xh_mlp_num_params(const xh_mlp* net, int64_t* n)
{
    if(!net || !n) return XH_EINVAL;
    *n = net->nparams;
    return XH_OK;
}

extern "C" int xh_mlp_get_params(xh_ctx* ctx, const xh_mlp* net, float* host)
{
    return guard(ctx, [&] {
        check(hipMemcpyAsync(host, net->params, size_t(net->nparams) * sizeof(float), hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync");
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

extern "C" int xh_mlp_set_params(xh_ctx* ctx, xh_mlp* net, const float* host)
{
    return guard(ctx, [&] {
        check(hipMemcpyAsync(net->params, host, size_t(net->nparams) * sizeof(float), hipMemcpyHostToDevice, ctx->stream), "hipMemcpyAsync");
        check(hipMemsetAsync(net->m, 0, size_t(net->nparams) * sizeof(float), ctx->stream), "hipMemsetAsync");
        check(hipMemsetAsync(net->v, 0, size_t(net->nparams) * sizeof(float), ctx->stream), "hipMemsetAsync");
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

extern "C" int xh_mlp_forward(xh_ctx* ctx, xh_mlp* net, const float* X, int64_t B, float* Y)
{
    return guard(ctx, [&] {
        if(B <= 0) return;
        ensure_capacity(ctx, net, B);
        forward(ctx, net, X, B, Y);
    });
}

extern "C" int xh_mlp_grad(xh_ctx* ctx, xh_mlp* net, const float* X, const float* T, int64_t B, float* grad, float* loss)
{
    return guard(ctx, [&] {
        if(B <= 0) throw xh::Error(XH_EINVAL, "mlp: empty batch");
        ensure_capacity(ctx, net, B);
        forward(ctx, net, X, B, net->y);
        const Layout L = layout(net->nin, net->width);
        const int w = net->width, nin = net->nin;
        const unsigned nb = blocks_for(B);
        hipLaunchKernelGGL(k_mse, dim3(nb), dim3(MT), 0, ctx->stream, net->y, T, B, net->dy, net->loss_part);
        hipLaunchKernelGGL(k_back_out, dim3(blocks_for(std::max<int64_t>(B * w, w))), dim3(MT), 0, ctx->stream, net->h2, net->dy, net->params + L.w3, w, B,
                           grad + L.w3, grad + L.b3, net->d2);
        hipLaunchKernelGGL(k_back_weights, dim3(blocks_for(int64_t(w) * w)), dim3(MT), 0, ctx->stream, net->h1, w, net->d2, w, B, grad + L.W2, grad + L.b2);
        hipLaunchKernelGGL(k_back_input, dim3(blocks_for(B * w)), dim3(MT), 0, ctx->stream, net->d2, net->params + L.W2, w, w, net->h1, B, net->d1);
        hipLaunchKernelGGL(k_back_weights, dim3(blocks_for(int64_t(nin) * w)), dim3(MT), 0, ctx->stream, X, nin, net->d1, w, B, grad + L.W1, grad + L.b1);
        xh::launch_check("mlp backward");
        if(loss) {
            std::vector<float> part(nb);
            check(hipMemcpyAsync(part.data(), net->loss_part, nb * sizeof(float), hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync");
            check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
            double acc = 0.0;
            for(float x : part) acc += x;
            *loss = static_cast<float>(acc / double(B));
        }
    });
}

extern "C" int xh_mlp_adam_step(xh_ctx* ctx, xh_mlp* net, const float* grad, float lr, int64_t t)
{
    return guard(ctx, [&] {
        if(t < 1) throw xh::Error(XH_EINVAL, "mlp: Adam step counter starts at 1");
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, float(t)), c2 = 1.f - std::pow(b2, float(t));
        hipLaunchKernelGGL(k_adam, dim3(blocks_for(net->nparams)), dim3(MT), 0, ctx->stream, net->params, grad, net->m, net->v, net->nparams, lr, b1, b2,
                           1e-8f, c1, c2);
        xh::launch_check("k_adam");
    });
}
