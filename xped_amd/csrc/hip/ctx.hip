// Context, buffers, events and the elementwise / reduction kernels of the device layer.
#include <algorithm>
#include <cfloat>
#include <cstring>

#include "device.hpp"

using xh::check;
using xh::guard;

extern "C" int xh_device_count(int* count)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if(e != hipSuccess) n = 0;
    *count = n;
    return n > 0 ? XH_OK : XH_ENODEV;
}

extern "C" int xh_create(int device, xh_ctx** out)
{
    *out = nullptr;
    int n = 0;
    if(hipGetDeviceCount(&n) != hipSuccess || n <= device) return XH_ENODEV;
    auto* ctx = new xh_ctx();
    int rc = guard(ctx, [&] {
        ctx->device = device;
        check(hipSetDevice(device), "hipSetDevice");
        hipDeviceProp_t prop;
        check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
        ctx->num_cu = prop.multiProcessorCount;
        // default priority stream, not synchronising with the null stream
        check(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
        for(auto& ev : ctx->events) check(hipEventCreate(&ev), "hipEventCreate");
        hipMemPool_t pool;
        check(hipDeviceGetDefaultMemPool(&pool, device), "hipDeviceGetDefaultMemPool");
        uint64_t thr = UINT64_MAX;
        check(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr), "hipMemPoolSetAttribute");
        check(hipHostMalloc(reinterpret_cast<void**>(&ctx->host_scalar), 64 * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
        check(hipMalloc(reinterpret_cast<void**>(&ctx->dev_scalar), (xh::RED_PARTS + 64) * sizeof(double)), "hipMalloc");
        check(hipMemset(ctx->dev_scalar + xh::RED_PARTS, 0, 64 * sizeof(double)), "hipMemset");  // normalisation status
        xh::gemm_scratch_init(ctx);
    });
    if(rc != XH_OK) {
        delete ctx;
        return rc;
    }
    *out = ctx;
    return XH_OK;
}

extern "C" int xh_destroy(xh_ctx* ctx)
{
    if(!ctx) return XH_OK;
    hipSetDevice(ctx->device);
    if(ctx->stream) hipStreamSynchronize(ctx->stream);
    xh::collect_staging(ctx, true);
    xh::pinned_destroy_all(ctx);
    xh::dcache_release(ctx);
    if(ctx->stream) hipStreamSynchronize(ctx->stream);
    for(auto& ev : ctx->events)
        if(ev) hipEventDestroy(ev);
    for(auto& ev : ctx->gemm_events) hipEventDestroy(ev);
    for(auto& ev : ctx->perm_events) hipEventDestroy(ev);
    if(ctx->host_scalar) hipHostFree(ctx->host_scalar);
    if(ctx->dev_scalar) hipFree(ctx->dev_scalar);
    if(ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
    return XH_OK;
}

extern "C" const char* xh_last_error(const xh_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

extern "C" int xh_alloc(xh_ctx* ctx, size_t bytes, void** dptr)
{
    return guard(ctx, [&] {
        *dptr = nullptr;
        if(bytes == 0) return;
        check(xh::dmalloc(ctx, dptr, bytes), "hipMallocAsync");
    });
}

extern "C" int xh_free(xh_ctx* ctx, void* dptr)
{
    return guard(ctx, [&] {
        if(dptr) check(xh::dfree(ctx, dptr), "hipFreeAsync");
    });
}

namespace {
// size class of a request: 1 KiB, or the request rounded up to a quarter of its power of two
// (2^e < b <= 2^(e+1) -> a multiple of 2^(e-2): 1.25, 1.5, 1.75 or 2 x 2^e)
size_t dclass(size_t b)
{
    if(b <= 1024) return 1024;
    const int e = 63 - __builtin_clzll(static_cast<unsigned long long>(b - 1));
    const size_t q = size_t(1) << (e - 2);
    return (b + q - 1) / q * q;
}
}  // namespace

hipError_t xh::dmalloc(xh_ctx* ctx, void** p, size_t bytes)
{
    *p = nullptr;
    if(xh_graph_capturing(ctx)) return hipMallocAsync(p, bytes, ctx->stream);  // owned by the capture
    const size_t c = dclass(bytes);
    auto it = ctx->dblock_idle.find(c);
    if(it != ctx->dblock_idle.end() && !it->second.empty()) {
        *p = it->second.back();
        it->second.pop_back();
        ctx->dcache_bytes -= c;
        ctx->dblock_live.emplace(*p, c);
        ++ctx->dcache_hits;
        return hipSuccess;
    }
    hipError_t e = hipMallocAsync(p, c, ctx->stream);
    if((e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) && ctx->dcache_bytes > 0) {
        (void)hipGetLastError();
        dcache_release(ctx);
        e = hipMallocAsync(p, c, ctx->stream);
    }
    if(e != hipSuccess) {
        *p = nullptr;
        return e;
    }
    ctx->dblock_live.emplace(*p, c);
    ++ctx->dcache_misses;
    return hipSuccess;
}

hipError_t xh::dfree(xh_ctx* ctx, void* p)
{
    if(!p) return hipSuccess;
    auto it = ctx->dblock_live.find(p);
    if(it == ctx->dblock_live.end()) return hipFreeAsync(p, ctx->stream);  // not from the cache
    const size_t c = it->second;
    ctx->dblock_live.erase(it);
    if(xh_graph_capturing(ctx) || ctx->dcache_bytes + c > ctx->dcache_limit) return hipFreeAsync(p, ctx->stream);
    ctx->dblock_idle[c].push_back(p);
    ctx->dcache_bytes += c;
    return hipSuccess;
}

void xh::dcache_release(xh_ctx* ctx)
{
    for(auto& kv : ctx->dblock_idle)
        for(void* p : kv.second) hipFreeAsync(p, ctx->stream);
    ctx->dblock_idle.clear();
    ctx->dcache_bytes = 0;
}

extern "C" int xh_set_device_cache(xh_ctx* ctx, size_t bytes)
{
    return guard(ctx, [&] {
        ctx->dcache_limit = bytes;
        if(ctx->dcache_bytes > bytes) xh::dcache_release(ctx);
    });
}

extern "C" int xh_device_cache_stats(xh_ctx* ctx, int64_t* hits, int64_t* misses, size_t* idle_bytes)
{
    if(!ctx) return XH_EINVAL;
    if(hits) *hits = ctx->dcache_hits;
    if(misses) *misses = ctx->dcache_misses;
    if(idle_bytes) *idle_bytes = ctx->dcache_bytes;
    return XH_OK;
}

extern "C" int xh_upload(xh_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    return guard(ctx, [&] {
        if(bytes == 0) return;
        check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream), "hipMemcpyAsync H2D");
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

extern "C" int xh_download(xh_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    return guard(ctx, [&] {
        if(bytes == 0) return;
        check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync D2H");
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

extern "C" int xh_copy(xh_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    return guard(ctx, [&] {
        if(bytes == 0) return;
        check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream), "hipMemcpyAsync D2D");
    });
}

extern "C" int xh_synchronize(xh_ctx* ctx)
{
    return guard(ctx, [&] {
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        xh::collect_staging(ctx, true);
    });
}

extern "C" void* xh_stream(xh_ctx* ctx) { return ctx ? reinterpret_cast<void*>(ctx->stream) : nullptr; }

struct xh_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<xh::PlanPin> pins;  // plans whose device tables the graph's kernels read
};

extern "C" int xh_graph_capturing(xh_ctx* ctx)
{
    if(!ctx) return 0;
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if(hipStreamIsCapturing(ctx->stream, &st) != hipSuccess) return 0;
    return st == hipStreamCaptureStatusActive ? 1 : 0;
}

void xh::refuse_under_capture(xh_ctx* ctx, const char* what)
{
    if(xh_graph_capturing(ctx)) {
        ctx->capture_failed = true;
        throw xh::Error(XH_EINVAL, std::string(what) + " cannot run under graph capture (build plans before capturing)");
    }
}

void xh::pin_for_capture(xh_ctx* ctx, PlanLife* plan, void (*release)(xh_ctx*, PlanLife*))
{
    if(!plan || !xh_graph_capturing(ctx)) return;
    ++plan->pins;
    ctx->capture_pins.push_back(PlanPin{plan, release});
}

void xh::unpin(xh_ctx* ctx, const PlanPin& p)
{
    if(--p.plan->pins == 0 && p.plan->dead) p.release(ctx, p.plan);
}

namespace {
void unpin_all(xh_ctx* ctx, std::vector<xh::PlanPin>& pins)
{
    for(const auto& p : pins) xh::unpin(ctx, p);
    pins.clear();
}
}  // namespace

extern "C" int xh_graph_begin(xh_ctx* ctx)
{
    return guard(ctx, [&] {
        if(xh_graph_capturing(ctx)) throw xh::Error(XH_EINVAL, "graph capture already active on this context");
        ctx->capture_failed = false;
        unpin_all(ctx, ctx->capture_pins);
        check(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    });
}

extern "C" int xh_graph_end(xh_ctx* ctx, xh_graph** out)
{
    if(out) *out = nullptr;
    return guard(ctx, [&] {
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        const bool refused = ctx->capture_failed;
        ctx->capture_failed = false;
        std::vector<xh::PlanPin> pins;
        pins.swap(ctx->capture_pins);
        if(e != hipSuccess || !g || refused) {
            unpin_all(ctx, pins);
            if(g) hipGraphDestroy(g);
            (void)hipGetLastError();
            if(refused) throw xh::Error(XH_EINVAL, "graph capture discarded: a captured call could not run under capture");
            throw xh::Error(XH_EHIP, std::string("graph capture failed: ") + hipGetErrorString(e));
        }
        auto* gr = new xh_graph();
        gr->graph = g;
        const hipError_t ei = hipGraphInstantiate(&gr->exec, g, nullptr, nullptr, 0);
        if(ei != hipSuccess) {
            unpin_all(ctx, pins);
            hipGraphDestroy(g);
            delete gr;
            throw xh::Error(XH_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        }
        if(!out) {
            unpin_all(ctx, pins);
            hipGraphExecDestroy(gr->exec);
            hipGraphDestroy(gr->graph);
            delete gr;
            throw xh::Error(XH_EINVAL, "xh_graph_end: null output");
        }
        gr->pins = std::move(pins);
        *out = gr;
    });
}

extern "C" int xh_graph_launch(xh_ctx* ctx, const xh_graph* g)
{
    return guard(ctx, [&] {
        if(!g || !g->exec) throw xh::Error(XH_EINVAL, "null graph");
        check(hipGraphLaunch(g->exec, ctx->stream), "hipGraphLaunch");
    });
}

extern "C" int xh_graph_destroy(xh_ctx* ctx, xh_graph* g)
{
    if(!g) return XH_OK;
    return guard(ctx, [&] {
        if(g->exec) check(hipGraphExecDestroy(g->exec), "hipGraphExecDestroy");
        if(g->graph) check(hipGraphDestroy(g->graph), "hipGraphDestroy");
        g->exec = nullptr;
        g->graph = nullptr;
        unpin_all(ctx, g->pins);  // frees the plans whose destroy was deferred for this graph
        delete g;
    });
}

extern "C" int xh_gemm_launch_log(xh_ctx* ctx, int64_t launch, double* ms, int32_t* trans, int32_t* mnk, int64_t cap, int64_t* nitems)
{
    return guard(ctx, [&] {
        const size_t n = ctx->gemm_events_used / 2;
        if(launch < 0 || size_t(launch) >= n || size_t(launch) >= ctx->gemm_launch_items.size())
            throw xh::Error(XH_EINVAL, "xh_gemm_launch_log: no such timed launch");
        check(hipEventSynchronize(ctx->gemm_events[2 * launch + 1]), "hipEventSynchronize");
        float t = 0.f;
        check(hipEventElapsedTime(&t, ctx->gemm_events[2 * launch], ctx->gemm_events[2 * launch + 1]), "hipEventElapsedTime");
        if(ms) *ms = t;
        const auto& v = ctx->gemm_launch_items[launch];
        const int64_t ni = (int64_t(v.size()) - 3) / 3;
        if(trans) {
            trans[0] = v[1];
            trans[1] = v[2];
            trans[2] = v[0];
        }
        if(nitems) *nitems = ni;
        if(mnk)
            for(int64_t i = 0; i < std::min(ni, cap); ++i)
                for(int j = 0; j < 3; ++j) mnk[3 * i + j] = v[3 + 3 * i + j];
    });
}

extern "C" int xh_set_gemm_tag(xh_ctx* ctx, int tag)
{
    return guard(ctx, [&] { ctx->gemm_tag = tag; });
}

extern "C" int xh_event_record(xh_ctx* ctx, int slot)
{
    return guard(ctx, [&] {
        if(slot < 0 || slot >= 16) throw xh::Error(XH_EINVAL, "event slot out of range");
        check(hipEventRecord(ctx->events[slot], ctx->stream), "hipEventRecord");
    });
}

extern "C" int xh_gemm_timing(xh_ctx* ctx, int enable, int reset, double* kernel_ms, double* flops, int64_t* launches)
{
    return guard(ctx, [&] {
        double ms = 0.0, fl = 0.0;
        const size_t n = ctx->gemm_events_used / 2;
        if(n > 0) check(hipEventSynchronize(ctx->gemm_events[ctx->gemm_events_used - 1]), "hipEventSynchronize");
        for(size_t i = 0; i < n; ++i) {
            float t = 0.f;
            check(hipEventElapsedTime(&t, ctx->gemm_events[2 * i], ctx->gemm_events[2 * i + 1]), "hipEventElapsedTime");
            ms += t;
            fl += ctx->gemm_event_flops[i];
        }
        if(kernel_ms) *kernel_ms = ms;
        if(flops) *flops = fl;
        if(launches) *launches = static_cast<int64_t>(n);
        if(reset) {
            ctx->gemm_events_used = 0;
            ctx->gemm_event_flops.clear();
            ctx->gemm_launch_items.clear();
        }
        if(enable >= 0) ctx->gemm_timing = enable != 0;
    });
}

extern "C" int xh_permute_timing(xh_ctx* ctx, int enable, int reset, double* kernel_ms, double* bytes, int64_t* launches)
{
    return guard(ctx, [&] {
        double ms = 0.0, b = 0.0;
        const size_t n = ctx->perm_events_used / 2;
        if(n > 0) check(hipEventSynchronize(ctx->perm_events[ctx->perm_events_used - 1]), "hipEventSynchronize");
        for(size_t i = 0; i < n; ++i) {
            float t = 0.f;
            check(hipEventElapsedTime(&t, ctx->perm_events[2 * i], ctx->perm_events[2 * i + 1]), "hipEventElapsedTime");
            ms += t;
            b += ctx->perm_event_bytes[i];
        }
        if(kernel_ms) *kernel_ms = ms;
        if(bytes) *bytes = b;
        if(launches) *launches = static_cast<int64_t>(n);
        if(reset) {
            ctx->perm_events_used = 0;
            ctx->perm_event_bytes.clear();
        }
        if(enable >= 0) ctx->perm_timing = enable != 0;
    });
}

namespace {
__global__ void k_trace_marker() {}
}  // namespace

extern "C" int xh_trace_marker(xh_ctx* ctx, int id)
{
    return guard(ctx, [&] {
        if(id < 1 || id > 4096) throw xh::Error(XH_EINVAL, "marker id must be in 1..4096");
        hipLaunchKernelGGL(k_trace_marker, dim3(static_cast<unsigned>(id)), dim3(64), 0, ctx->stream);
        xh::launch_check("k_trace_marker");
    });
}

extern "C" int xh_event_elapsed_ms(xh_ctx* ctx, int s0, int s1, float* ms)
{
    return guard(ctx, [&] {
        if(s0 < 0 || s0 >= 16 || s1 < 0 || s1 >= 16) throw xh::Error(XH_EINVAL, "event slot out of range");
        check(hipEventSynchronize(ctx->events[s1]), "hipEventSynchronize");
        check(hipEventElapsedTime(ms, ctx->events[s0], ctx->events[s1]), "hipEventElapsedTime");
    });
}

// ------------------------------------------------------------------ elementwise kernels

namespace {

constexpr int EW_THREADS = 256;

inline unsigned ew_grid(size_t n, int per_thread = 4)
{
    size_t g = (n + static_cast<size_t>(EW_THREADS) * per_thread - 1) / (static_cast<size_t>(EW_THREADS) * per_thread);
    return static_cast<unsigned>(std::max<size_t>(1, std::min<size_t>(g, 65535u * 16u)));
}

__global__ void k_fill(double* __restrict__ x, size_t n, double v)
{
    for(size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) x[i] = v;
}

// ranges r = blockIdx.y of (offset, length): from the kernel arguments (up to FR_ARGS ranges) or
// from a device table (off, len pairs)
constexpr int FR_ARGS = 8;
struct FillRanges {
    int64_t off[FR_ARGS], len[FR_ARGS];
};
__global__ void k_fill_ranges(double* __restrict__ x, FillRanges fr, const int64_t* __restrict__ table, double v)
{
    const int r = blockIdx.y;
    const int64_t off = table ? table[2 * r] : fr.off[r], len = table ? table[2 * r + 1] : fr.len[r];
    double* p = x + off;
    for(int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < len; e += int64_t(gridDim.x) * blockDim.x) p[e] = v;
}

__global__ void k_scale(double* __restrict__ x, size_t n, double a)
{
    for(size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) x[i] = a * x[i];
}

// y = x / m with m = *res, the max-abs of x folded on the device.  A vanished or non-finite m
// copies x unchanged and is counted in status[0] (the first such m kept in status[1]).
__global__ void k_scale_inv_dev(const double* __restrict__ x, double* __restrict__ y, size_t n, const double* __restrict__ res,
                                double* __restrict__ status)
{
    const double m = *res;
    const bool ok = m > 0.0 && m <= DBL_MAX;  // NaN fails both
    if(!ok && blockIdx.x == 0 && threadIdx.x == 0) {
        if(status[0] == 0.0) status[1] = m;
        status[0] += 1.0;
    }
    const double a = ok ? 1.0 / m : 1.0;
    for(size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) y[i] = a * x[i];
}

// y = a x + b y; with b == 0, y is not read (it may be uninitialised, and 0 * NaN is NaN), and
// the result is the plain product a x, as the reference's scalar multiplication computes it
__global__ void k_axpby(double* __restrict__ y, const double* __restrict__ x, size_t n, double a, double b)
{
    if(b == 0.0) {
        for(size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) y[i] = __dmul_rn(a, x[i]);
        return;
    }
    for(size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        y[i] = __dadd_rn(__dmul_rn(a, x[i]), __dmul_rn(b, y[i]));
}

// Two-stage reductions: per-block partials into scratch, then a single block folds them.
// max that propagates NaN (fmax would drop it), so that maxNorm of non-finite data is non-finite
__device__ inline double nmax(double a, double b) { return (a > b || a != a) ? a : b; }
__device__ inline double wave_max(double v)
{
    for(int o = 32; o > 0; o >>= 1) v = nmax(v, __shfl_xor(v, o));
    return v;
}
__device__ inline double wave_sum(double v)
{
    for(int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__global__ void k_maxabs(const double* __restrict__ x, size_t n, double* __restrict__ part)
{
    double m = 0.0;
    for(size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) m = nmax(m, fabs(x[i]));
    m = wave_max(m);
    __shared__ double s[EW_THREADS / 64];
    if((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
    __syncthreads();
    if(threadIdx.x == 0) {
        double r = s[0];
        for(int w = 1; w < EW_THREADS / 64; ++w) r = nmax(r, s[w]);
        part[blockIdx.x] = r;
    }
}

__global__ void k_dot(const double* __restrict__ x, const double* __restrict__ y, size_t n, double* __restrict__ part)
{
    double acc = 0.0;
    for(size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) acc = fma(x[i], y[i], acc);
    acc = wave_sum(acc);
    __shared__ double s[EW_THREADS / 64];
    if((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
    __syncthreads();
    if(threadIdx.x == 0) {
        double r = 0;
        for(int w = 0; w < EW_THREADS / 64; ++w) r += s[w];
        part[blockIdx.x] = r;
    }
}

__global__ void k_fold(const double* __restrict__ part, int n, double* __restrict__ out, int is_max)
{
    double acc = 0.0;
    for(int i = threadIdx.x; i < n; i += blockDim.x) acc = is_max ? nmax(acc, part[i]) : acc + part[i];
    acc = is_max ? wave_max(acc) : wave_sum(acc);
    __shared__ double s[EW_THREADS / 64];
    if((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
    __syncthreads();
    if(threadIdx.x == 0) {
        double r = s[0];
        for(int w = 1; w < EW_THREADS / 64; ++w) r = is_max ? nmax(r, s[w]) : r + s[w];
        *out = r;
    }
}

__global__ void k_block_trace(const int64_t* __restrict__ off, const int32_t* __restrict__ m, const double* __restrict__ x, double* __restrict__ out)
{
    const int g = blockIdx.x;
    const int64_t o = off[g];
    const int mm = m[g];
    double acc = 0.0;
    for(int i = threadIdx.x; i < mm; i += blockDim.x) acc += x[o + int64_t(i) * mm + i];
    acc = wave_sum(acc);
    __shared__ double s[EW_THREADS / 64];
    if((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
    __syncthreads();
    if(threadIdx.x == 0) {
        double r = 0;
        for(int w = 0; w < EW_THREADS / 64; ++w) r += s[w];
        out[g] = r;
    }
}

__global__ void k_scale_rects(double* __restrict__ x, const xh_rect* __restrict__ rects)
{
    const xh_rect r = rects[blockIdx.x];
    const int64_t n = int64_t(r.rows) * r.cols;
    for(int64_t e = threadIdx.x; e < n; e += blockDim.x) {
        int64_t i = e % r.rows, j = e / r.rows;
        double* p = x + r.off + i + j * r.ld;
        *p = r.factor * *p;
    }
}

__global__ void k_diag_map(const int64_t* __restrict__ off, const int32_t* __restrict__ m, double* __restrict__ x, int mode)
{
    const int g = blockIdx.x;
    const int64_t o = off[g];
    const int mm = m[g];
    const int f = mode & 3;
    for(int i = threadIdx.x; i < mm; i += blockDim.x) {
        double* p = x + o + int64_t(i) * mm + i;
        double v = *p;
        if(f == 0)
            v = sqrt(v);
        else if(f == 1)
            v = 1.0 / v;
        else
            v = 1.0 / sqrt(v);
        *p = v;
    }
    if(mode & 4) {  // diagonal matrix result: off-diagonal entries become zero
        const int64_t n2 = int64_t(mm) * mm;
        for(int64_t e = threadIdx.x; e < n2; e += blockDim.x)
            if(e % mm != e / mm) x[o + e] = 0.0;
    }
}

// exp(factor * M) of square blocks of order <= XH_MEXP_MAX (TensorBase::mexp, TensorBase.cpp:180-183:
// Eigen's matrix exponential of each block, MatrixInterface_Eigen_impl.hpp:177-180).  Scaling and
// squaring: B = factor * M / 2^s with ||B||_1 <= 1/2, exp(B) by its Taylor sum to degree
// MEXP_DEGREE (remainder below 0.5^19 / 19! ~ 1.6e-23 relative to ||exp(B)||), then s squarings.
// One workgroup per block; the three m x m matrices live in LDS and each thread keeps its share of
// a product in registers until the barrier that frees the operand it overwrites.
constexpr int MEXP_THREADS = 256;
constexpr int MEXP_PER_THREAD = XH_MEXP_MAX * XH_MEXP_MAX / MEXP_THREADS;
constexpr int MEXP_DEGREE = 18;
static_assert(XH_MEXP_MAX * XH_MEXP_MAX % MEXP_THREADS == 0, "mexp tile split");

__device__ inline void mexp_product(const double* X, const double* Y, int m, double* acc)
{
    const int n2 = m * m;
#pragma unroll
    for(int r = 0; r < MEXP_PER_THREAD; ++r) {
        const int e = threadIdx.x + r * MEXP_THREADS;
        double v = 0.0;
        if(e < n2) {
            const int i = e % m, j = e / m;
            for(int l = 0; l < m; ++l) v += X[i + l * m] * Y[l + j * m];
        }
        acc[r] = v;
    }
}

__global__ __launch_bounds__(MEXP_THREADS) void k_block_mexp(const xh_mexp_item* __restrict__ items, const double* src, double* dst,
                                                            double factor)
{
    __shared__ double B[XH_MEXP_MAX * XH_MEXP_MAX], T[XH_MEXP_MAX * XH_MEXP_MAX], S[XH_MEXP_MAX * XH_MEXP_MAX];
    __shared__ double colsum[XH_MEXP_MAX];
    __shared__ int nsq;
    const xh_mexp_item it = items[blockIdx.x];
    const int m = it.m, n2 = m * m;
    for(int e = threadIdx.x; e < n2; e += MEXP_THREADS) B[e] = factor * src[it.src_off + e];
    __syncthreads();
    if(threadIdx.x < m) {
        double c = 0.0;
        for(int i = 0; i < m; ++i) c += fabs(B[i + threadIdx.x * m]);
        colsum[threadIdx.x] = c;
    }
    __syncthreads();
    if(threadIdx.x == 0) {
        double nrm = 0.0;
        for(int j = 0; j < m; ++j) nrm = fmax(nrm, colsum[j]);
        int s = 0;
        if(nrm > 0.5 && isfinite(nrm)) {
            int ex = 0;
            frexp(nrm, &ex);  // nrm = f 2^ex, f in [1/2, 1): nrm / 2^(ex + 1) < 1/2
            s = ex + 1;
        }
        nsq = s;
    }
    __syncthreads();
    const double down = ldexp(1.0, -nsq);
    for(int e = threadIdx.x; e < n2; e += MEXP_THREADS) {
        const double b = B[e] * down;
        B[e] = b;
        T[e] = b;
        S[e] = b + ((e % m) == (e / m) ? 1.0 : 0.0);
    }
    __syncthreads();
    double acc[MEXP_PER_THREAD];
    for(int k = 2; k <= MEXP_DEGREE; ++k) {
        mexp_product(T, B, m, acc);  // T_k = T_{k-1} B / k
        __syncthreads();
        const double inv_k = 1.0 / k;
#pragma unroll
        for(int r = 0; r < MEXP_PER_THREAD; ++r) {
            const int e = threadIdx.x + r * MEXP_THREADS;
            if(e < n2) {
                const double t = acc[r] * inv_k;
                T[e] = t;
                S[e] += t;
            }
        }
        __syncthreads();
    }
    for(int q = 0; q < nsq; ++q) {
        mexp_product(S, S, m, acc);
        __syncthreads();
#pragma unroll
        for(int r = 0; r < MEXP_PER_THREAD; ++r) {
            const int e = threadIdx.x + r * MEXP_THREADS;
            if(e < n2) S[e] = acc[r];
        }
        __syncthreads();
    }
    for(int e = threadIdx.x; e < n2; e += MEXP_THREADS) dst[it.dst_off + e] = S[e];
}

// Reverse passes of the diagonal maps (ADTensor.hpp:560-590), diagonal entries only, as the
// reference's diagonal expressions are diagonal matrices: mode 0 (diag_sqrt, y = sqrt(x)):
// dx_ii += dy_ii * 0.5 / y_ii;  mode 1 (diag_inv, y = 1/x): dx_ii += -dy_ii * y_ii^2.
__global__ void k_diag_bwd(const int64_t* __restrict__ off, const int32_t* __restrict__ m, const double* __restrict__ y, const double* __restrict__ dy,
                           double* __restrict__ dx, int mode)
{
    const int g = blockIdx.x;
    const int64_t o = off[g];
    const int mm = m[g];
    for(int i = threadIdx.x; i < mm; i += blockDim.x) {
        const int64_t e = o + int64_t(i) * mm + i;
        const double yv = y[e], g1 = dy[e];
        double d;
        if(mode == 0)
            d = g1 * (1.0 / yv) * 0.5;  // y = sqrt(x)
        else if(mode == 1)
            d = g1 * (-(yv * yv));  // y = 1 / x
        else
            d = g1 * (-0.5 * yv * yv * yv);  // y = 1 / sqrt(x)
        dx[e] += d;
    }
}

// x_g += w_g * identity for every block g (rows x cols, ld rows): the trace reverse pass with
// WeightedIdentity (ADTensor.hpp:533-545, Tensor.hpp:238-246).
__global__ void k_add_identity(const int64_t* __restrict__ off, const int32_t* __restrict__ rows, const int32_t* __restrict__ cols,
                               const double* __restrict__ w, double* __restrict__ x)
{
    const int g = blockIdx.x;
    const int r = rows[g], n = min(r, cols[g]);
    for(int i = threadIdx.x; i < n; i += blockDim.x) x[off[g] + int64_t(i) * r + i] += w[g];
}

// Largest |x_i| and the first index attaining it (NaN counts as largest), two passes.
struct MaxIdx {
    double v;
    int64_t i;
};
__device__ inline MaxIdx better(MaxIdx a, MaxIdx b)
{
    const bool an = isnan(a.v), bn = isnan(b.v);
    if(an != bn) return an ? a : b;
    if(a.v != b.v && !an) return a.v > b.v ? a : b;
    return a.i <= b.i ? a : b;
}
__global__ void k_maxabs_idx(const double* __restrict__ x, size_t n, MaxIdx* __restrict__ part)
{
    MaxIdx m{-1.0, INT64_MAX};
    for(size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) m = better(m, MaxIdx{fabs(x[i]), int64_t(i)});
    __shared__ MaxIdx s[EW_THREADS];
    s[threadIdx.x] = m;
    __syncthreads();
    for(int h = EW_THREADS / 2; h > 0; h >>= 1) {
        if(int(threadIdx.x) < h) s[threadIdx.x] = better(s[threadIdx.x], s[threadIdx.x + h]);
        __syncthreads();
    }
    if(threadIdx.x == 0) part[blockIdx.x] = s[0];
}
__global__ void k_fold_idx(MaxIdx* __restrict__ part, int nb)
{
    __shared__ MaxIdx s[EW_THREADS];
    MaxIdx m{-1.0, INT64_MAX};
    for(int i = threadIdx.x; i < nb; i += blockDim.x) m = better(m, part[i]);
    s[threadIdx.x] = m;
    __syncthreads();
    for(int h = EW_THREADS / 2; h > 0; h >>= 1) {
        if(int(threadIdx.x) < h) s[threadIdx.x] = better(s[threadIdx.x], s[threadIdx.x + h]);
        __syncthreads();
    }
    if(threadIdx.x == 0) part[nb] = s[0];
}

__global__ void k_add_at(double* __restrict__ x, int64_t i, double v)
{
    if(threadIdx.x == 0) x[i] += v;
}

// Rectangles in tiles of 32 x 32, blockIdx.y the rectangle, the tiles spread over blockIdx.x:
// straight copies read and write along the rectangle's columns; transposed ones stage the tile in
// LDS so that both the reads (along the source's columns) and the writes are contiguous.
constexpr int CR_TILE = 32, CR_THREADS = 256;
template <bool ACC>
__global__ __launch_bounds__(CR_THREADS) void k_copy_rects(const double* __restrict__ src, double* __restrict__ dst, const xh_copy_rect* __restrict__ rects)
{
    __shared__ double tile[CR_TILE][CR_TILE + 1];
    const xh_copy_rect r = rects[blockIdx.y];
    const int ntr = (r.rows + CR_TILE - 1) / CR_TILE, ntc = (r.cols + CR_TILE - 1) / CR_TILE;
    const int tx = threadIdx.x % CR_TILE, ty = threadIdx.x / CR_TILE;
    constexpr int STEP = CR_THREADS / CR_TILE;
    auto put = [&](int i, int j, double v) {
        v = r.factor == 1.0 ? v : r.factor * v;
        double* p = dst + r.dst_off + i + int64_t(j) * r.dst_ld;
        if constexpr(ACC)
            *p += v;
        else
            *p = v;
    };
    for(int t = blockIdx.x; t < ntr * ntc; t += gridDim.x) {
        const int i0 = (t % ntr) * CR_TILE, j0 = (t / ntr) * CR_TILE;
        if(!r.transpose) {
            for(int q = ty; q < CR_TILE; q += STEP) {
                const int i = i0 + tx, j = j0 + q;
                if(i < r.rows && j < r.cols) put(i, j, src[r.src_off + i + int64_t(j) * r.src_ld]);
            }
        } else {
            // dst (i, j) = src (j, i); the source is cols x rows
            __syncthreads();  // the previous tile's reads of the LDS tile are done
            for(int q = ty; q < CR_TILE; q += STEP) {
                const int j = j0 + tx, i = i0 + q;
                tile[q][tx] = (i < r.rows && j < r.cols) ? src[r.src_off + j + int64_t(i) * r.src_ld] : 0.0;
            }
            __syncthreads();
            for(int q = ty; q < CR_TILE; q += STEP) {
                const int i = i0 + tx, j = j0 + q;
                if(i < r.rows && j < r.cols) put(i, j, tile[tx][q]);
            }
        }
    }
}

// Counter-based Gaussian fill: element i takes two 53-bit uniforms from splitmix64 (Steele, Lea &
// Flood, OOPSLA 2014) of (seed, i) and one Box-Muller normal (Box & Muller, Ann. Math. Stat. 29
// (1958) 610).  Deterministic for a given seed and independent of the launch shape.
__device__ inline uint64_t splitmix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__global__ void k_fill_gaussian(double* __restrict__ x, size_t n, uint64_t seed)
{
    const uint64_t key = splitmix64(seed);
    for(size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
        const uint64_t a = splitmix64(key ^ (2 * uint64_t(i))), b = splitmix64(key ^ (2 * uint64_t(i) + 1));
        const double u1 = (double((a >> 11) + 1)) * 0x1.0p-53;  // (0, 1]
        const double u2 = double(b >> 11) * 0x1.0p-53;          // [0, 1)
        x[i] = sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
    }
}

// Per range g: out[g] = sum of squares of x[off[g] .. off[g] + len[g]).
__global__ void k_range_sumsq(const int64_t* __restrict__ off, const int64_t* __restrict__ len, const double* __restrict__ x, double* __restrict__ out)
{
    const int g = blockIdx.x;
    const double* p = x + off[g];
    const int64_t n = len[g];
    double acc = 0.0;
    for(int64_t i = threadIdx.x; i < n; i += blockDim.x) acc = fma(p[i], p[i], acc);
    acc = wave_sum(acc);
    __shared__ double s[EW_THREADS / 64];
    if((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
    __syncthreads();
    if(threadIdx.x == 0) {
        double r = 0;
        for(int w = 0; w < EW_THREADS / 64; ++w) r += s[w];
        out[g] = r;
    }
}

}  // namespace

extern "C" int xh_fill(xh_ctx* ctx, double* dst, size_t n, double value)
{
    return guard(ctx, [&] {
        if(n == 0) return;
        hipLaunchKernelGGL(k_fill, dim3(ew_grid(n)), dim3(EW_THREADS), 0, ctx->stream, dst, n, value);
        xh::launch_check("k_fill");
    });
}

extern "C" int xh_fill_ranges(xh_ctx* ctx, double* dst, size_t n, const int64_t* ranges, int64_t nranges, double value)
{
    return guard(ctx, [&] {
        if(nranges < 0) throw xh::Error(XH_EINVAL, "fill ranges: negative count");
        int64_t maxlen = 0;
        for(int64_t r = 0; r < nranges; ++r) {
            const int64_t off = ranges[2 * r], len = ranges[2 * r + 1];
            if(off < 0 || len < 0 || static_cast<uint64_t>(off) + static_cast<uint64_t>(len) > n)
                throw xh::Error(XH_EINVAL, "fill ranges: a range lies outside the buffer");
            maxlen = std::max(maxlen, len);
        }
        if(maxlen == 0) return;
        const dim3 block(EW_THREADS);
        const unsigned gx = std::min(ew_grid(static_cast<size_t>(maxlen)), 1024u);
        for(int64_t r0 = 0; r0 < nranges;) {
            if(nranges - r0 <= FR_ARGS) {
                FillRanges fr{};
                const int nr = static_cast<int>(nranges - r0);
                for(int i = 0; i < nr; ++i) {
                    fr.off[i] = ranges[2 * (r0 + i)];
                    fr.len[i] = ranges[2 * (r0 + i) + 1];
                }
                hipLaunchKernelGGL(k_fill_ranges, dim3(gx, nr), block, 0, ctx->stream, dst, fr, static_cast<const int64_t*>(nullptr), value);
                xh::launch_check("k_fill_ranges");
                break;
            }
            // more ranges than fit in the arguments: a table of at most 65535 per launch
            const int64_t nr = std::min<int64_t>(nranges - r0, 65535);
            std::vector<int64_t> t(ranges + 2 * r0, ranges + 2 * (r0 + nr));
            xh::PackedUpload up;
            const size_t o = up.add(t);
            up.upload(ctx);
            hipLaunchKernelGGL(k_fill_ranges, dim3(gx, static_cast<unsigned>(nr)), block, 0, ctx->stream, dst, FillRanges{}, up.ptr<const int64_t>(o), value);
            xh::launch_check("k_fill_ranges");
            up.release_deferred(ctx);
            r0 += nr;
        }
    });
}

extern "C" int xh_scale(xh_ctx* ctx, double* x, size_t n, double a)
{
    return guard(ctx, [&] {
        if(n == 0) return;
        hipLaunchKernelGGL(k_scale, dim3(ew_grid(n)), dim3(EW_THREADS), 0, ctx->stream, x, n, a);
        xh::launch_check("k_scale");
    });
}

extern "C" int xh_axpby(xh_ctx* ctx, double* y, const double* x, size_t n, double a, double b)
{
    return guard(ctx, [&] {
        if(n == 0) return;
        hipLaunchKernelGGL(k_axpby, dim3(ew_grid(n)), dim3(EW_THREADS), 0, ctx->stream, y, x, n, a, b);
        xh::launch_check("k_axpby");
    });
}

static void reduce(xh_ctx* ctx, const double* x, const double* y, size_t n, double* out, bool is_max)
{
    if(n == 0) {
        *out = 0.0;
        return;
    }
    // the maximum does not depend on the order of its partials, so it spreads over up to
    // RED_PARTS workgroups (the whole chip); the sum keeps 32 partials in a fixed order
    const size_t want = (n + EW_THREADS - 1) / EW_THREADS;
    const int nb = static_cast<int>(std::min<size_t>(is_max ? xh::RED_PARTS : 32, want));
    double* res = ctx->dev_scalar + xh::RED_PARTS;
    if(is_max)
        hipLaunchKernelGGL(k_maxabs, dim3(nb), dim3(EW_THREADS), 0, ctx->stream, x, n, ctx->dev_scalar);
    else
        hipLaunchKernelGGL(k_dot, dim3(nb), dim3(EW_THREADS), 0, ctx->stream, x, y, n, ctx->dev_scalar);
    xh::launch_check("reduce");
    hipLaunchKernelGGL(k_fold, dim3(1), dim3(EW_THREADS), 0, ctx->stream, ctx->dev_scalar, nb, res, is_max ? 1 : 0);
    xh::launch_check("fold");
    check(hipMemcpyAsync(ctx->host_scalar, res, sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync");
    check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    *out = ctx->host_scalar[0];
}

extern "C" int xh_max_abs(xh_ctx* ctx, const double* x, size_t n, double* out)
{
    return guard(ctx, [&] { reduce(ctx, x, nullptr, n, out, true); });
}

extern "C" int xh_normalize_max_abs(xh_ctx* ctx, const double* x, double* y, size_t n)
{
    // asynchronous: the norm never leaves the device
    return guard(ctx, [&] {
        if(n == 0) return;
        const size_t want = (n + EW_THREADS - 1) / EW_THREADS;
        const int nb = static_cast<int>(std::min<size_t>(xh::RED_PARTS, want));
        double* res = ctx->dev_scalar + xh::RED_PARTS;
        hipLaunchKernelGGL(k_maxabs, dim3(nb), dim3(EW_THREADS), 0, ctx->stream, x, n, ctx->dev_scalar);
        xh::launch_check("k_maxabs");
        hipLaunchKernelGGL(k_fold, dim3(1), dim3(EW_THREADS), 0, ctx->stream, ctx->dev_scalar, nb, res, 1);
        xh::launch_check("fold");
        hipLaunchKernelGGL(k_scale_inv_dev, dim3(ew_grid(n)), dim3(EW_THREADS), 0, ctx->stream, x, y, n, res, res + 2);
        xh::launch_check("k_scale_inv_dev");
    });
}

extern "C" int xh_normalize_status(xh_ctx* ctx, int64_t* count, double* first)
{
    return guard(ctx, [&] {
        xh::refuse_under_capture(ctx, "normalize status");
        double* st = ctx->dev_scalar + xh::RED_PARTS + 2;
        check(hipMemcpyAsync(ctx->host_scalar + 8, st, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync");
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        *count = static_cast<int64_t>(ctx->host_scalar[8]);
        *first = ctx->host_scalar[9];
        if(*count) check(hipMemsetAsync(st, 0, 2 * sizeof(double), ctx->stream), "hipMemsetAsync");
    });
}

extern "C" int xh_dot(xh_ctx* ctx, const double* x, const double* y, size_t n, double* out)
{
    return guard(ctx, [&] { reduce(ctx, x, y, n, out, false); });
}

extern "C" int xh_block_trace(xh_ctx* ctx, const int64_t* off, const int32_t* m, int64_t nblocks, const double* x, double* out)
{
    return guard(ctx, [&] {
        if(nblocks <= 0) return;
        std::vector<int64_t> o(off, off + nblocks);
        std::vector<int32_t> mm(m, m + nblocks);
        int64_t* d_off = xh::dev_upload(ctx, o);
        int32_t* d_m = xh::dev_upload(ctx, mm);
        double* d_out = nullptr;
        check(xh::dmalloc(ctx, reinterpret_cast<void**>(&d_out), nblocks * sizeof(double)), "hipMallocAsync");
        hipLaunchKernelGGL(k_block_trace, dim3(static_cast<unsigned>(nblocks)), dim3(EW_THREADS), 0, ctx->stream, d_off, d_m, x, d_out);
        xh::launch_check("k_block_trace");
        check(hipMemcpyAsync(out, d_out, nblocks * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync");
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        xh::dev_free(ctx, d_off);
        xh::dev_free(ctx, d_m);
        xh::dev_free(ctx, d_out);
    });
}

extern "C" int xh_scale_rects(xh_ctx* ctx, double* x, const xh_rect* rects, int64_t nrects)
{
    return guard(ctx, [&] {
        if(nrects <= 0) return;
        std::vector<xh_rect> r(rects, rects + nrects);
        xh::PackedUpload up;
        const size_t o = up.add(r);
        up.upload(ctx);
        hipLaunchKernelGGL(k_scale_rects, dim3(static_cast<unsigned>(nrects)), dim3(EW_THREADS), 0, ctx->stream, x, up.ptr<const xh_rect>(o));
        xh::launch_check("k_scale_rects");
        up.release_deferred(ctx);
    });
}

extern "C" int xh_block_diag_map(xh_ctx* ctx, const int64_t* off, const int32_t* m, int64_t nblocks, double* x, int mode)
{
    return guard(ctx, [&] {
        if(nblocks <= 0) return;
        if((mode & 3) > 2 || mode < 0 || mode > 6) throw xh::Error(XH_EINVAL, "diag map mode must be 0, 1 or 2, optionally + 4");
        std::vector<int64_t> o(off, off + nblocks);
        std::vector<int32_t> mm(m, m + nblocks);
        xh::PackedUpload up;
        const size_t oo = up.add(o), om = up.add(mm);
        up.upload(ctx);
        hipLaunchKernelGGL(k_diag_map, dim3(static_cast<unsigned>(nblocks)), dim3(EW_THREADS), 0, ctx->stream, up.ptr<const int64_t>(oo),
                           up.ptr<const int32_t>(om), x, mode);
        xh::launch_check("k_diag_map");
        up.release_deferred(ctx);
    });
}

extern "C" int xh_block_mexp(xh_ctx* ctx, const xh_mexp_item* items, int64_t nitems, const double* src, double* dst, double factor)
{
    return guard(ctx, [&] {
        if(nitems <= 0) return;
        std::vector<xh_mexp_item> v;
        for(int64_t i = 0; i < nitems; ++i) {
            const xh_mexp_item& it = items[i];
            if(it.m < 0 || it.m > XH_MEXP_MAX) throw xh::Error(XH_EINVAL, "mexp: block order must be in [0, 64]");
            if(it.src_off < 0 || it.dst_off < 0) throw xh::Error(XH_EINVAL, "mexp: negative offset");
            if(it.m > 0) v.push_back(it);
        }
        if(v.empty()) return;
        xh::PackedUpload up;
        const size_t o = up.add(v);
        up.upload(ctx);
        hipLaunchKernelGGL(k_block_mexp, dim3(static_cast<unsigned>(v.size())), dim3(MEXP_THREADS), 0, ctx->stream, up.ptr<const xh_mexp_item>(o), src, dst,
                           factor);
        xh::launch_check("k_block_mexp");
        up.release_deferred(ctx);
    });
}

extern "C" int xh_block_diag_bwd(xh_ctx* ctx, const int64_t* off, const int32_t* m, int64_t nblocks, const double* y, const double* dy, double* dx, int mode)
{
    return guard(ctx, [&] {
        if(nblocks <= 0) return;
        if(mode < 0 || mode > 2) throw xh::Error(XH_EINVAL, "diag adjoint mode must be 0 (sqrt), 1 (inverse) or 2 (inverse sqrt)");
        std::vector<int64_t> o(off, off + nblocks);
        std::vector<int32_t> mm(m, m + nblocks);
        xh::PackedUpload up;
        const size_t oo = up.add(o), om = up.add(mm);
        up.upload(ctx);
        hipLaunchKernelGGL(k_diag_bwd, dim3(static_cast<unsigned>(nblocks)), dim3(EW_THREADS), 0, ctx->stream, up.ptr<const int64_t>(oo),
                           up.ptr<const int32_t>(om), y, dy, dx, mode);
        xh::launch_check("k_diag_bwd");
        up.release_deferred(ctx);
    });
}

extern "C" int xh_block_add_identity(xh_ctx* ctx, const int64_t* off, const int32_t* rows, const int32_t* cols, const double* w, int64_t nblocks, double* x)
{
    return guard(ctx, [&] {
        if(nblocks <= 0) return;
        std::vector<int64_t> o(off, off + nblocks);
        std::vector<int32_t> r(rows, rows + nblocks), c(cols, cols + nblocks);
        std::vector<double> ww(w, w + nblocks);
        xh::PackedUpload up;
        const size_t oo = up.add(o), orr = up.add(r), oc = up.add(c), ow = up.add(ww);
        up.upload(ctx);
        hipLaunchKernelGGL(k_add_identity, dim3(static_cast<unsigned>(nblocks)), dim3(EW_THREADS), 0, ctx->stream, up.ptr<const int64_t>(oo),
                           up.ptr<const int32_t>(orr), up.ptr<const int32_t>(oc), up.ptr<const double>(ow), x);
        xh::launch_check("k_add_identity");
        up.release_deferred(ctx);
    });
}

extern "C" int xh_max_abs_index(xh_ctx* ctx, const double* x, size_t n, double* value, int64_t* index)
{
    return guard(ctx, [&] {
        *value = 0.0;
        *index = -1;
        if(n == 0) return;
        const unsigned nb = std::min(ew_grid(n), 1024u);
        MaxIdx* part = nullptr;
        check(xh::dmalloc(ctx, reinterpret_cast<void**>(&part), (nb + 1) * sizeof(MaxIdx)), "hipMallocAsync");
        hipLaunchKernelGGL(k_maxabs_idx, dim3(nb), dim3(EW_THREADS), 0, ctx->stream, x, n, part);
        xh::launch_check("k_maxabs_idx");
        hipLaunchKernelGGL(k_fold_idx, dim3(1), dim3(EW_THREADS), 0, ctx->stream, part, static_cast<int>(nb));
        xh::launch_check("k_fold_idx");
        MaxIdx r;
        check(hipMemcpyAsync(&r, part + nb, sizeof(MaxIdx), hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync");
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        xh::dev_free(ctx, part);
        *value = r.v;
        *index = r.i;
    });
}

extern "C" int xh_add_at(xh_ctx* ctx, double* x, int64_t index, double v)
{
    return guard(ctx, [&] {
        if(index < 0) throw xh::Error(XH_EINVAL, "add_at: negative index");
        hipLaunchKernelGGL(k_add_at, dim3(1), dim3(64), 0, ctx->stream, x, index, v);
        xh::launch_check("k_add_at");
    });
}

static int copy_rects_impl(bool acc, xh_ctx* ctx, const double* src, double* dst, const xh_copy_rect* rects, int64_t nrects)
{
    return guard(ctx, [&] {
        if(nrects <= 0) return;
        if(nrects > 65535) throw xh::Error(XH_EINVAL, "copy_rects: at most 65535 rectangles per call");
        std::vector<xh_copy_rect> r(rects, rects + nrects);
        int64_t tiles = 0;
        for(const auto& x : r) {
            if(x.rows < 0 || x.cols < 0 || x.src_ld < 0 || x.dst_ld < 0) throw xh::Error(XH_EINVAL, "copy_rects: negative extent or leading dimension");
            tiles = std::max(tiles, int64_t((x.rows + CR_TILE - 1) / CR_TILE) * ((x.cols + CR_TILE - 1) / CR_TILE));
        }
        if(tiles == 0) return;
        xh::PackedUpload up;
        const size_t o = up.add(r);
        up.upload(ctx);
        const dim3 grid(static_cast<unsigned>(std::min<int64_t>(tiles, 64)), static_cast<unsigned>(nrects));
        hipLaunchKernelGGL(acc ? k_copy_rects<true> : k_copy_rects<false>, grid, dim3(CR_THREADS), 0, ctx->stream, src, dst, up.ptr<const xh_copy_rect>(o));
        xh::launch_check("k_copy_rects");
        up.release_deferred(ctx);
    });
}

extern "C" int xh_copy_rects(xh_ctx* ctx, const double* src, double* dst, const xh_copy_rect* rects, int64_t nrects)
{
    return copy_rects_impl(false, ctx, src, dst, rects, nrects);
}

extern "C" int xh_add_rects(xh_ctx* ctx, const double* src, double* dst, const xh_copy_rect* rects, int64_t nrects)
{
    return copy_rects_impl(true, ctx, src, dst, rects, nrects);
}

extern "C" int xh_fill_gaussian(xh_ctx* ctx, double* dst, size_t n, uint64_t seed)
{
    return guard(ctx, [&] {
        if(n == 0) return;
        hipLaunchKernelGGL(k_fill_gaussian, dim3(ew_grid(n)), dim3(EW_THREADS), 0, ctx->stream, dst, n, seed);
        xh::launch_check("k_fill_gaussian");
    });
}

extern "C" int xh_range_sumsq(xh_ctx* ctx, const int64_t* off, const int64_t* len, int64_t nranges, const double* x, double* out)
{
    return guard(ctx, [&] {
        if(nranges <= 0) return;
        std::vector<int64_t> v(off, off + nranges);
        v.insert(v.end(), len, len + nranges);
        int64_t* d = xh::dev_upload(ctx, v);
        double* d_out = nullptr;
        check(xh::dmalloc(ctx, reinterpret_cast<void**>(&d_out), size_t(nranges) * sizeof(double)), "hipMallocAsync");
        hipLaunchKernelGGL(k_range_sumsq, dim3(static_cast<unsigned>(nranges)), dim3(EW_THREADS), 0, ctx->stream, d, d + nranges, x, d_out);
        xh::launch_check("k_range_sumsq");
        check(hipMemcpyAsync(out, d_out, size_t(nranges) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync");
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        xh::dev_free(ctx, d);
        xh::dev_free(ctx, d_out);
    });
}
